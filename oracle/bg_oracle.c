/*
 * bg_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see bg_oracle.h).
 *
 * Restates, function by function, the reference's single-threaded path. Every floating-point
 * expression keeps the reference's C++ types (float BGFLOAT, double literals, float overloads of
 * <cmath> under `using namespace std`, -std=c++98 pow(float,float)), because bit-exactness against
 * the compiled reference is how this oracle is pinned (tests/test_oracle_vs_reference.py).
 * Built with -O2 -ffp-contract=off on x86-64 SSE, like the reference's g++ -O2 build.
 */
#include "bg_oracle.h"
#include "../include/bg_detmath.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ULONG_MAX64 0xFFFFFFFFFFFFFFFFull
enum { INH = 1, EXC = 2 };
enum { II = 0, IE = 1, EI = 2, EE = 3 };

/* ------------------------------------------------------------------ MTRand (MT19937) ------ */
/* RNG/MersenneTwister.cpp:66-291, restated from Matsumoto & Nishimura's MT19937 definition. */
typedef struct { uint32_t s[624]; int left; int next; } mt_t;

static void mt_initialize(mt_t *m, uint32_t seed)
{   /* RNG/MersenneTwister.cpp:260-275 */
    m->s[0] = seed;
    for (int i = 1; i < 624; i++) m->s[i] = 1812433253u * (m->s[i - 1] ^ (m->s[i - 1] >> 30)) + (uint32_t)i;
}
static inline uint32_t mt_twist(uint32_t mm, uint32_t s0, uint32_t s1)
{   /* RNG/MersenneTwister.h:141-144 */
    return mm ^ (((s0 & 0x80000000u) | (s1 & 0x7fffffffu)) >> 1) ^ ((uint32_t)(-(int32_t)(s1 & 1u)) & 0x9908b0dfu);
}
static void mt_reload(mt_t *m)
{   /* RNG/MersenneTwister.cpp:278-291 */
    int i = 0;
    for (; i < 624 - 397; i++) m->s[i] = mt_twist(m->s[i + 397], m->s[i], m->s[i + 1]);
    for (; i < 623; i++) m->s[i] = mt_twist(m->s[i + 397 - 624], m->s[i], m->s[i + 1]);
    m->s[623] = mt_twist(m->s[623 + 397 - 624], m->s[623], m->s[0]);
    m->left = 624;
    m->next = 0;
}
static void mt_seed(mt_t *m, uint32_t seed) { mt_initialize(m, seed); mt_reload(m); }
static uint32_t mt_randint(mt_t *m)
{   /* RNG/MersenneTwister.cpp:93-107 */
    if (m->left == 0) mt_reload(m);
    --m->left;
    uint32_t s1 = m->s[m->next++];
    s1 ^= (s1 >> 11);
    s1 ^= (s1 << 7) & 0x9d2c5680u;
    s1 ^= (s1 << 15) & 0xefc60000u;
    return s1 ^ (s1 >> 18);
}
/* MTRand::rand (RNG/MersenneTwister.cpp:75-76): float(randInt) * (1.0/4294967295.0) -> float */
static float mt_rand(mt_t *m) { return (float)((double)(float)mt_randint(m) * (1.0 / 4294967295.0)); }
/* MTRand::inRange (RNG/MersenneTwister.cpp:128-134) */
static float mt_inrange(mt_t *m, float lo, float hi)
{
    float val = mt_rand(m);
    float range = hi - lo;
    val *= range;
    val += lo;
    return val;
}

/* Norm (RNG/Norm.cpp:59-86, Norm.h:69-70): Marsaglia polar pairs, mu 0, sigma 1. */
typedef struct { mt_t mt; int odd; float x2; } norm_t;
static void norm_seed(norm_t *g, uint32_t seed) { mt_seed(&g->mt, seed); g->odd = 1; g->x2 = 0.0f; }
static float norm_next(norm_t *g)
{
    const float mu = 0.0f, sigma = 1.0f;
    if ((g->odd = !g->odd)) return mu + sigma * g->x2;
    float U1, U2, V1, V2, W;
    do {
        U1 = mt_rand(&g->mt);
        U2 = mt_rand(&g->mt);
        V1 = 2 * U1 - 1;
        V2 = 2 * U2 - 1;
        W = V1 * V1 + V2 * V2;
    } while (W > 1);
    float Y = sqrtf(-2 * logf(W) / W);
    float X1 = V1 * Y;
    g->x2 = V2 * Y;
    return mu + sigma * X1;
}

/* ------------------------------------------------------------------ network state --------- */
struct bgo_net {
    bgo_params p;
    int N, M, W32, R, max_spikes;
    uint64_t step;
    int epoch;                  /* completed growth updates */
    char err[256];
    int has_error;
    mt_t rng;                   /* global `rng(1)` (Utils/Global.cpp:79) */
    norm_t norm;                /* rgNormrnd[0] (Core/SingleThreadedSpikingModel.cpp:30) */
    float (*f_expf)(float);
    float (*f_acosf)(float);
    float (*f_sinf)(float);
    float (*f_powf)(float, float);
    double (*f_exp)(double);
    /* layout */
    float *xloc, *yloc;
    int32_t *ntype;
    uint8_t *starter;
    int32_t n_starters;
    /* neurons */
    float *Vm, *Vthresh, *Vreset, *Vrest, *Vinit, *sp, *I0, *Inoise, *Iinject, *C1, *C2, *Trefract, *Tau, *Rm, *Cm;
    float *Aconst, *Bconst, *Cconst, *Dconst, *u, *C3;
    int32_t *refr;
    uint8_t *hasFired;
    int32_t *spikeCount, *spikeCountOffset;
    uint64_t *hist;
    /* synapses: slot arrays of N*M */
    uint8_t *in_use;
    int32_t *src, *dst, *type, *total_delay, *delayIdx, *delayIdxPost, *total_delayPost;
    float *W, *psr, *decay, *tau, *r, *us, *U, *D, *F;
    uint64_t *lastSpike;
    uint32_t *ring, *ringPost;
    float *Apos, *Aneg, *STDPgap, *tauspost, *tauspre, *taupos, *tauneg, *Wex, *mupos, *muneg;
    uint8_t *useFD;
    uint32_t *syn_counts;
    int64_t total;
    /* index map (Core/SynapseIndexMap.h) */
    uint32_t *in_map, *in_begin, *in_count, *out_map, *out_begin, *out_count;
    int has_map;
    /* growth */
    float *radii, *rates;
    float *radiiHistory, *ratesHistory, *burst, *spikesHist;
    int64_t n_burst, n_spikesHist;
    int32_t *scratch_slot;      /* N, -1 */
    /* stimulus input: one DS input synapse per neuron (Inputs/SInputPoisson.cpp:124-146) */
    uint8_t *in_mask;
    int32_t *in_nisi, *in_type, *in_total_delay, *in_delayIdx, *in_shift;
    float *in_W, *in_psr, *in_decay, *in_r, *in_u, *in_U, *in_D, *in_F, *in_values;
    uint64_t *in_lastSpike;
    uint32_t *in_ring;
    float in_lambda;            /* mean inter-spike interval in ms */
    int32_t in_dur, in_cycle, in_steps_in_cycle;
};

static int input_construct(bgo_net *n);
static void input_init(bgo_net *n);
static void input_stimulus(bgo_net *n);

static int fail(bgo_net *n, const char *msg)
{
    if (!n->has_error) snprintf(n->err, sizeof n->err, "%s", msg);
    n->has_error = 1;
    return -1;
}

static float own_expf(float x) { return bg_expf(x); }
static float own_acosf(float x) { return bg_acosf(x); }
static float own_sinf(float x) { return bg_sinf(x); }
static float own_powf(float x, float y) { return bg_powf(x, y); }
static double own_exp(double x) { return bg_exp(x); }
static float lib_expf(float x) { return expf(x); }
static float lib_acosf(float x) { return acosf(x); }
static float lib_sinf(float x) { return sinf(x); }
static float lib_powf(float x, float y) { return powf(x, y); }
static double lib_exp(double x) { return exp(x); }

#define ALLOC(ptr, cnt) do { (ptr) = calloc(((size_t)(cnt) != 0) ? (size_t)(cnt) : 1, sizeof(*(ptr))); if (!(ptr)) goto oom; } while (0)

static int syn_sign(int t) { return (t == II || t == IE) ? -1 : 1; }  /* Synapses/AllSynapses.cpp:385-400 */
static int syn_type(const bgo_net *n, int s, int d)
{   /* Layouts/Layout.cpp:129-141 */
    int a = n->ntype[s], b = n->ntype[d];
    if (a == INH && b == INH) return II;
    if (a == INH && b == EXC) return IE;
    if (a == EXC && b == INH) return EI;
    return EE;
}

/* ------------------------------------------------------------------ setup ----------------- */
static void layout_setup(bgo_net *n)
{
    int N = n->N;
    for (int i = 0; i < N; i++) {   /* Layouts/Layout.cpp:155-157 (grid) */
        n->xloc[i] = (float)(i % n->p.width);
        n->yloc[i] = (float)(i / n->p.width);
        n->ntype[i] = EXC;          /* Layouts/Layout.cpp:104-106 */
        n->starter[i] = 0;
    }
    if (n->p.layout_kind == BGO_LAYOUT_FIXED) {
        for (int i = 0; i < n->p.n_inh; i++) n->ntype[n->p.inh_list[i]] = INH;   /* FixedLayout.cpp:258-261 */
        for (int i = 0; i < n->p.n_active; i++) n->starter[n->p.active_list[i]] = 1; /* :275-278 */
        n->n_starters = n->p.n_active;
    } else {
        /* DynamicLayout::generateNeuronTypeMap (Layouts/DynamicLayout.cpp:98-130) */
        int num_exc = (int)((double)(n->p.frac_exc * N) + 0.5);
        int num_inh = N - num_exc;
        int *rg = malloc(sizeof(int) * (num_inh > 0 ? num_inh : 1));
        for (int i = 0; i < num_inh; i++) rg[i] = i;
        for (int i = num_inh; i < N; i++) {
            int j = (int)(mt_rand(&n->rng) * N);
            if (j < num_inh) rg[j] = i;
        }
        for (int i = 0; i < num_inh; i++) n->ntype[rg[i]] = INH;
        free(rg);
        /* DynamicLayout::initStarterMap (Layouts/DynamicLayout.cpp:139-165) */
        uint32_t num_st = (uint32_t)((double)(n->p.frac_starter * N) + 0.5);
        uint32_t alloc = 0;
        while (alloc < num_st) {
            int i = (int)mt_inrange(&n->rng, 0.0f, (float)N);
            if (n->ntype[i] == EXC && !n->starter[i]) { n->starter[i] = 1; alloc++; }
        }
        n->n_starters = (int32_t)num_st;
    }
}

static void neurons_create(bgo_net *n)
{
    const bgo_params *p = &n->p;
    const float dt = p->dt;
    for (int i = 0; i < n->N; i++) {
        /* AllIFNeurons::setNeuronDefaults (Neurons/AllIFNeurons.cpp:459-471) */
        n->Cm[i] = (float)3e-8; n->Rm[i] = (float)1e6; n->Vthresh[i] = (float)-0.04;
        n->Vrest[i] = 0.0f; n->Vreset[i] = (float)-0.06; n->Vinit[i] = (float)-0.06;
        n->Trefract[i] = (float)3e-3; n->Inoise[i] = 0.0f; n->Iinject[i] = 0.0f;
        n->Tau[i] = (float)(3e-8 * 1e6);
        if (p->neuron_kind == BGO_IZH) n->Trefract[i] = 0.0f;   /* Neurons/AllIZHNeurons.cpp:456 */
        /* AllIFNeurons::createNeuron (Neurons/AllIFNeurons.cpp:398-441) */
        n->Iinject[i] = mt_inrange(&n->rng, p->Iinject[0], p->Iinject[1]);
        n->Inoise[i] = mt_inrange(&n->rng, p->Inoise[0], p->Inoise[1]);
        n->Vthresh[i] = mt_inrange(&n->rng, p->Vthresh[0], p->Vthresh[1]);
        n->Vrest[i] = mt_inrange(&n->rng, p->Vresting[0], p->Vresting[1]);
        n->Vreset[i] = mt_inrange(&n->rng, p->Vreset[0], p->Vreset[1]);
        n->Vinit[i] = mt_inrange(&n->rng, p->Vinit[0], p->Vinit[1]);
        n->Vm[i] = n->Vinit[i];
        /* initNeuronConstsFromParamValues (Neurons/AllIFNeurons.cpp:479-503); host libm, as the
         * reference computes these once on the host. */
        if (n->Tau[i] > 0) {
            n->C1[i] = expf(-dt / n->Tau[i]);
            n->C2[i] = n->Rm[i] * (1 - n->C1[i]);
        } else {
            n->C1[i] = 0.0f;
            n->C2[i] = n->Rm[i];
        }
        n->I0[i] = n->Iinject[i] + n->Vrest[i] / n->Rm[i];
        if (p->neuron_kind == BGO_IZH) n->C3[i] = dt * 1000;   /* Neurons/AllIZHNeurons.cpp:474-475 */
        for (int j = 0; j < n->max_spikes; j++) n->hist[(size_t)i * n->max_spikes + j] = ULONG_MAX64;
        n->Trefract[i] = (n->ntype[i] == INH) ? (float)2.0e-3 : (float)3.0e-3;
        if (n->starter[i]) {
            n->Vthresh[i] = mt_inrange(&n->rng, p->starter_Vthresh[0], p->starter_Vthresh[1]);
            n->Vreset[i] = mt_inrange(&n->rng, p->starter_Vreset[0], p->starter_Vreset[1]);
            n->Trefract[i] = (float)3.0e-3;
        }
        if (p->neuron_kind == BGO_IZH) {   /* AllIZHNeurons::createNeuron (:413-433) */
            if (n->ntype[i] == EXC) {
                n->Aconst[i] = mt_inrange(&n->rng, p->excA[0], p->excA[1]);
                n->Bconst[i] = mt_inrange(&n->rng, p->excB[0], p->excB[1]);
                n->Cconst[i] = mt_inrange(&n->rng, p->excC[0], p->excC[1]);
                n->Dconst[i] = mt_inrange(&n->rng, p->excD[0], p->excD[1]);
            } else {
                n->Aconst[i] = mt_inrange(&n->rng, p->inhA[0], p->inhA[1]);
                n->Bconst[i] = mt_inrange(&n->rng, p->inhB[0], p->inhB[1]);
                n->Cconst[i] = mt_inrange(&n->rng, p->inhC[0], p->inhC[1]);
                n->Dconst[i] = mt_inrange(&n->rng, p->inhD[0], p->inhD[1]);
            }
            n->u[i] = 0.0f;
        }
    }
}

void bgo_destroy(bgo_net *n)
{
    if (!n) return;
    void *ptrs[] = {n->xloc, n->yloc, n->ntype, n->starter, n->Vm, n->Vthresh, n->Vreset, n->Vrest, n->Vinit,
                    n->sp, n->I0, n->Inoise, n->Iinject, n->C1, n->C2, n->Trefract, n->Tau, n->Rm, n->Cm,
                    n->Aconst, n->Bconst, n->Cconst, n->Dconst, n->u, n->C3, n->refr, n->hasFired,
                    n->spikeCount, n->spikeCountOffset, n->hist, n->in_use, n->src, n->dst, n->type,
                    n->total_delay, n->delayIdx, n->delayIdxPost, n->total_delayPost, n->W, n->psr, n->decay,
                    n->tau, n->r, n->us, n->U, n->D, n->F, n->lastSpike, n->ring, n->ringPost, n->Apos,
                    n->Aneg, n->STDPgap, n->tauspost, n->tauspre, n->taupos, n->tauneg, n->Wex, n->mupos,
                    n->muneg, n->useFD, n->syn_counts, n->in_map, n->in_begin, n->in_count, n->out_map,
                    n->out_begin, n->out_count, n->radii, n->rates, n->radiiHistory, n->ratesHistory,
                    n->burst, n->spikesHist, n->scratch_slot, n->in_mask, n->in_nisi, n->in_type,
                    n->in_total_delay, n->in_delayIdx, n->in_shift, n->in_W, n->in_psr, n->in_decay, n->in_r,
                    n->in_u, n->in_U, n->in_D, n->in_F, n->in_values, n->in_lastSpike, n->in_ring};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
    free(n);
}

static int conn_static_setup(bgo_net *n);

bgo_net *bgo_create(const bgo_params *p, char *err, int errlen)
{
    bgo_net *n = calloc(1, sizeof *n);
    if (!n) { if (err) snprintf(err, errlen, "out of memory"); return NULL; }
    n->p = *p;
    n->N = p->width * p->height;
    n->M = p->max_synapses_per_neuron;
    n->R = p->ring_bits ? p->ring_bits : 32;
    if (n->R % 32 || n->R > 1024 || n->N <= 0 || n->M < 0) {
        if (err) snprintf(err, errlen, "bad sizes (N=%d M=%d ring_bits=%d)", n->N, n->M, n->R);
        free(n);
        return NULL;
    }
    n->W32 = n->R / 32;
    n->max_spikes = (int)(p->epoch_duration * p->max_firing_rate);   /* AllIFNeurons.cpp:411 */
    if (p->math_mode) {
        n->f_expf = own_expf; n->f_acosf = own_acosf; n->f_sinf = own_sinf; n->f_powf = own_powf; n->f_exp = own_exp;
    } else {
        n->f_expf = lib_expf; n->f_acosf = lib_acosf; n->f_sinf = lib_sinf; n->f_powf = lib_powf; n->f_exp = lib_exp;
    }
    int N = n->N;
    size_t S = (size_t)N * (size_t)n->M;
    ALLOC(n->xloc, N); ALLOC(n->yloc, N); ALLOC(n->ntype, N); ALLOC(n->starter, N);
    ALLOC(n->Vm, N); ALLOC(n->Vthresh, N); ALLOC(n->Vreset, N); ALLOC(n->Vrest, N); ALLOC(n->Vinit, N);
    ALLOC(n->sp, N); ALLOC(n->I0, N); ALLOC(n->Inoise, N); ALLOC(n->Iinject, N); ALLOC(n->C1, N);
    ALLOC(n->C2, N); ALLOC(n->Trefract, N); ALLOC(n->Tau, N); ALLOC(n->Rm, N); ALLOC(n->Cm, N);
    ALLOC(n->Aconst, N); ALLOC(n->Bconst, N); ALLOC(n->Cconst, N); ALLOC(n->Dconst, N); ALLOC(n->u, N);
    ALLOC(n->C3, N); ALLOC(n->refr, N); ALLOC(n->hasFired, N); ALLOC(n->spikeCount, N);
    ALLOC(n->spikeCountOffset, N); ALLOC(n->hist, (size_t)N * (n->max_spikes > 0 ? n->max_spikes : 1));
    ALLOC(n->in_use, S); ALLOC(n->src, S); ALLOC(n->dst, S); ALLOC(n->type, S); ALLOC(n->total_delay, S);
    ALLOC(n->delayIdx, S); ALLOC(n->delayIdxPost, S); ALLOC(n->total_delayPost, S); ALLOC(n->W, S);
    ALLOC(n->psr, S); ALLOC(n->decay, S); ALLOC(n->tau, S); ALLOC(n->r, S); ALLOC(n->us, S); ALLOC(n->U, S);
    ALLOC(n->D, S); ALLOC(n->F, S); ALLOC(n->lastSpike, S); ALLOC(n->ring, S * n->W32);
    ALLOC(n->ringPost, S); ALLOC(n->Apos, S); ALLOC(n->Aneg, S); ALLOC(n->STDPgap, S); ALLOC(n->tauspost, S);
    ALLOC(n->tauspre, S); ALLOC(n->taupos, S); ALLOC(n->tauneg, S); ALLOC(n->Wex, S); ALLOC(n->mupos, S);
    ALLOC(n->muneg, S); ALLOC(n->useFD, S); ALLOC(n->syn_counts, N);
    ALLOC(n->in_map, S); ALLOC(n->in_begin, N); ALLOC(n->in_count, N); ALLOC(n->out_map, S);
    ALLOC(n->out_begin, N); ALLOC(n->out_count, N);
    ALLOC(n->radii, N); ALLOC(n->rates, N);
    int E = p->num_epochs > 0 ? p->num_epochs : 1;
    ALLOC(n->radiiHistory, (size_t)(E + 1) * N); ALLOC(n->ratesHistory, (size_t)(E + 1) * N);
    n->n_burst = (int64_t)(int)(p->epoch_duration * p->num_epochs);          /* XmlRecorder.cpp:14 */
    n->n_spikesHist = (int64_t)(int)(p->epoch_duration * p->num_epochs * 100); /* :15 */
    ALLOC(n->burst, n->n_burst); ALLOC(n->spikesHist, n->n_spikesHist);
    ALLOC(n->scratch_slot, N);
    for (int i = 0; i < N; i++) n->scratch_slot[i] = -1;
    ALLOC(n->in_mask, N); ALLOC(n->in_nisi, N); ALLOC(n->in_type, N); ALLOC(n->in_total_delay, N);
    ALLOC(n->in_delayIdx, N); ALLOC(n->in_shift, N); ALLOC(n->in_W, N); ALLOC(n->in_psr, N);
    ALLOC(n->in_decay, N); ALLOC(n->in_r, N); ALLOC(n->in_u, N); ALLOC(n->in_U, N); ALLOC(n->in_D, N);
    ALLOC(n->in_F, N); ALLOC(n->in_values, N); ALLOC(n->in_lastSpike, N); ALLOC(n->in_ring, N);

    /* The stimulus object is constructed before Simulator::setup (Core/BGDriver.cpp:76-87), so
     * SInputRegular's asynchronous shifts are the first draws of the global rng. */
    mt_seed(&n->rng, 1u);
    if (input_construct(n) < 0) {
        if (err) snprintf(err, errlen, "%s", n->err);
        bgo_destroy(n);
        return NULL;
    }
    /* Model::setupSim order (Core/Model.cpp:136-158) */
    layout_setup(n);
    neurons_create(n);
    for (int i = 0; i < N; i++) {   /* ConnGrowth::setupConnections (Connections/ConnGrowth.cpp:71-85) */
        n->radii[i] = p->start_radius;
        n->rates[i] = 0.0f;
        n->radiiHistory[i] = p->start_radius;   /* XmlGrowthRecorder::initDefaultValues (:27-37) */
        n->ratesHistory[i] = 0.0f;
    }
    /* m_conns->setupConnections, then createSynapseImap (Core/Model.cpp:154-157) */
    if (p->conn_kind == BGO_CONN_STATIC && (conn_static_setup(n) < 0 || bgo_rebuild_index_map(n) < 0)) {
        if (err) snprintf(err, errlen, "%s", n->err);
        bgo_destroy(n);
        return NULL;
    }
    norm_seed(&n->norm, p->seed);
    input_init(n);   /* Simulator::setup (Core/Simulator.cpp:40-43) */
    return n;
oom:
    if (err) snprintf(err, errlen, "out of memory");
    bgo_destroy(n);
    return NULL;
}

/* ------------------------------------------------------------------ synapses -------------- */
static inline int ring_test_clear(uint32_t *ring, int *idx, int R)
{   /* AllSpikingSynapses::isSpikeQueue (Synapses/AllSpikingSynapses.cpp:242-254) */
    int b = *idx;
    uint32_t m = 1u << (b & 31);
    int r = (ring[b >> 5] & m) != 0;
    ring[b >> 5] &= ~m;
    if (++(*idx) >= R) *idx = 0;
    return r;
}
static inline int ring_insert(uint32_t *ring, int idx, int delay, int R)
{   /* AllSpikingSynapses::preSpikeHit (Synapses/AllSpikingSynapses.cpp:261-279) */
    int b = idx + delay;
    if (b >= R) b -= R;
    uint32_t m = 1u << (b & 31);
    if (ring[b >> 5] & m) return -1;   /* assert(!(delay_queue & (0x1 << idx))) */
    ring[b >> 5] |= m;
    return 0;
}

/* createSynapse chain: AllSpikingSynapses.cpp:191-233, AllDSSynapses.cpp:160-218,
 * AllSTDPSynapses.cpp:124-138,252-273, AllDynamicSTDPSynapses.cpp:160-218. */
static int create_synapse(bgo_net *n, size_t i, int s, int d, int t)
{
    const float dt = n->p.dt;
    n->in_use[i] = 1;
    n->dst[i] = d;
    n->src[i] = s;
    n->W[i] = (float)(syn_sign(t) * 10.0e-9);
    n->type[i] = t;
    float tau, delay;
    static const float tau_tab[4] = {(float)6e-3, (float)6e-3, (float)3e-3, (float)3e-3};
    static const float delay_tab[4] = {(float)0.8e-3, (float)0.8e-3, (float)0.8e-3, (float)1.5e-3};
    tau = tau_tab[t];
    delay = n->p.delay_override[t] > 0 ? n->p.delay_override[t] : delay_tab[t];
    n->tau[i] = tau;
    n->total_delay[i] = (int)(delay / dt) + 1;
    if (n->total_delay[i] >= n->R) return fail(n, "synapse delay does not fit the delay ring");
    for (int w = 0; w < n->W32; w++) n->ring[i * n->W32 + w] = 0;
    n->delayIdx[i] = 0;
    n->psr[i] = 0.0f;
    n->decay[i] = expf(-dt / tau);   /* updateDecay (AllSpikingSynapses.cpp:344-354): host libm */
    int k = n->p.syn_kind;
    if (k == BGO_DS || k == BGO_DYNSTDP) {
        static const float U_tab[4] = {(float)0.32, (float)0.25, (float)0.05, (float)0.5};
        static const float D_tab[4] = {(float)0.144, (float)0.7, (float)0.125, (float)1.1};
        static const float F_tab[4] = {(float)0.06, (float)0.02, (float)1.2, (float)0.05};
        n->us[i] = (float)0.4;      /* DEFAULT_U */
        n->r[i] = 1.0f;
        n->lastSpike[i] = ULONG_MAX64;
        n->U[i] = U_tab[t]; n->D[i] = D_tab[t]; n->F[i] = F_tab[t];
    }
    if (k == BGO_STDP || k == BGO_DYNSTDP) {
        const bgo_params *p = &n->p;
        int o = p->stdp_override;
        n->Apos[i] = o ? p->stdp_Apos : 0.5f;
        n->Aneg[i] = o ? p->stdp_Aneg : -0.5f;
        n->STDPgap[i] = o ? p->stdp_gap : (float)2e-3;
        n->total_delayPost[i] = 0;
        n->tauspost[i] = o ? p->stdp_tauspost : (float)75e-3;
        n->tauspre[i] = o ? p->stdp_tauspre : (float)34e-3;
        n->taupos[i] = o ? p->stdp_taupos : (float)15e-3;
        n->tauneg[i] = o ? p->stdp_tauneg : (float)35e-3;
        n->Wex[i] = o ? p->stdp_Wex : 1.0f;
        n->mupos[i] = o ? p->stdp_mupos : 0.0f;
        n->muneg[i] = o ? p->stdp_muneg : 0.0f;
        n->useFD[i] = o ? (uint8_t)(p->stdp_use_froemke_dan != 0) : 1;
        n->ringPost[i] = 0;
        n->delayIdxPost[i] = 0;
    }
    return 0;
}

/* AllSynapses::addSynapse (Synapses/AllSynapses.cpp:357-377), first free slot of dest. */
static int64_t add_synapse(bgo_net *n, int s, int d, int t)
{
    if ((int)n->syn_counts[d] >= n->M) { fail(n, "destination neuron has no free synapse slot"); return -1; }
    size_t base = (size_t)d * n->M, i = base;
    for (int k = 0; k < n->M; k++) { i = base + k; if (!n->in_use[i]) break; }
    n->syn_counts[d]++;
    if (create_synapse(n, i, s, d, t) < 0) return -1;
    return (int64_t)i;
}

int64_t bgo_add_synapse(bgo_net *n, int32_t src, int32_t dest, float w)
{
    if (src < 0 || src >= n->N || dest < 0 || dest >= n->N) { fail(n, "neuron index out of range"); return -1; }
    int64_t i = add_synapse(n, src, dest, syn_type(n, src, dest));
    if (i >= 0) n->W[i] = w;
    return i;
}

/* ConnStatic::setupConnections (Connections/ConnStatic.cpp:33-88): for every source in ascending
 * order, the other neurons within threshConnsRadius sorted by distance (std::sort, equal distances
 * in the standard library's order), the first nConnsPerNeuron of them connected through addSynapse
 * and each weight drawn from the global rng right after its synapse: excitatory types from
 * [minExc, maxExc], inhibitory ones from [minInh, maxInh]. Distances are Layout::dist, the float
 * square root of the float squared distance (Layouts/Layout.cpp:54-69). A destination without a
 * free slot makes the reference write the weight through an unset index; here it fails, and a
 * destination that reaches maxSynapsesPerNeuron fails in createSynapseImap as the reference's
 * assert does (Synapses/AllSynapses.cpp:231). */
static int conn_static_setup(bgo_net *n)
{
    const int N = n->N;
    const bgo_params *p = &n->p;
    float *dist = malloc(sizeof(float) * (size_t)N);
    int32_t *dest = malloc(sizeof(int32_t) * (size_t)N);
    if (!dist || !dest) { free(dist); free(dest); return fail(n, "out of memory"); }
    int rc = 0;
    for (int s = 0; s < N && rc == 0; s++) {
        int c = 0;
        for (int d = 0; d < N; d++) {
            if (d == s) continue;
            float dx = n->xloc[s] - n->xloc[d], dy = n->yloc[s] - n->yloc[d];
            float d2 = dx * dx + dy * dy;
            float r = sqrtf(d2);
            if (r <= p->static_thresh) { dist[c] = r; dest[c] = d; c++; }
        }
        bgo_sort_by_dist(dist, dest, c);
        for (int i = 0; i < c && i < p->static_nconns; i++) {
            int d = dest[i], t = syn_type(n, s, d);
            int64_t slot = add_synapse(n, s, d, t);
            if (slot < 0) { rc = -1; break; }
            n->W[slot] = syn_sign(t) > 0 ? mt_inrange(&n->rng, p->static_exc_w[0], p->static_exc_w[1])
                                         : mt_inrange(&n->rng, p->static_inh_w[0], p->static_inh_w[1]);
        }
    }
    free(dist);
    free(dest);
    return rc;
}

/* AllSynapses::createSynapseImap (Synapses/AllSynapses.cpp:223-293). */
int bgo_rebuild_index_map(bgo_net *n)
{
    int N = n->N, M = n->M;
    int64_t total = 0;
    for (int i = 0; i < N; i++) {
        if (!((int)n->syn_counts[i] < M)) return fail(n, "synapse_counts reached maxSynapsesPerNeuron");
        total += n->syn_counts[i];
    }
    if (total == 0) return 0;   /* the reference keeps the previous map (AllSynapses.cpp:237-240) */
    int64_t k = 0;
    memset(n->out_count, 0, sizeof(uint32_t) * N);
    for (int i = 0; i < N; i++) {
        n->in_begin[i] = (uint32_t)k;
        uint32_t c = 0;
        for (int j = 0; j < M; j++) {
            size_t s = (size_t)i * M + j;
            if (n->in_use[s]) { n->in_map[k++] = (uint32_t)s; c++; n->out_count[n->src[s]]++; }
        }
        if (c != n->syn_counts[i]) return fail(n, "synapse_counts disagrees with in_use");
        n->in_count[i] = c;
    }
    uint64_t acc = 0;
    for (int i = 0; i < N; i++) { n->out_begin[i] = (uint32_t)acc; acc += n->out_count[i]; }
    uint32_t *fill = calloc(N, sizeof(uint32_t));
    for (int64_t q = 0; q < k; q++) {   /* stable by source, ascending iSyn */
        uint32_t s = n->in_map[q];
        int a = n->src[s];
        n->out_map[n->out_begin[a] + fill[a]++] = s;
    }
    free(fill);
    n->total = total;
    n->has_map = 1;
    return 0;
}

static inline uint64_t spike_hist(const bgo_net *n, int idx, int off)
{   /* AllSpikingNeurons::getSpikeHistory (Neurons/AllSpikingNeurons.cpp:182-188) */
    int ms = n->max_spikes;
    int k = (n->spikeCount[idx] + n->spikeCountOffset[idx] + ms + off) % ms;
    return n->hist[(size_t)idx * ms + k];
}

/* AllSTDPSynapses::stdpLearning (Synapses/AllSTDPSynapses.cpp:432-472) */
static void stdp_learning(bgo_net *n, size_t i, double delta, double epost, double epre)
{
    float STDPgap = n->STDPgap[i], muneg = n->muneg[i], mupos = n->mupos[i], tauneg = n->tauneg[i];
    float taupos = n->taupos[i], Aneg = n->Aneg[i], Apos = n->Apos[i], Wex = n->Wex[i];
    float W = n->W[i];
    float dw;
    if (delta < -STDPgap) {
        dw = (float)((double)(n->f_powf(W, muneg) * Aneg) * n->f_exp(delta / tauneg));
    } else if (delta > STDPgap) {
        dw = (float)((double)(n->f_powf(Wex - W, mupos) * Apos) * n->f_exp(-delta / taupos));
    } else {
        return;
    }
    W = (float)((double)W + epost * epre * (double)dw);
    if ((Wex < 0 && W > 0) || (Wex > 0 && W < 0)) W = 0;
    if (fabsf(W) > fabsf(Wex)) W = Wex;
    n->W[i] = W;
}

/* changePSR: AllSpikingSynapses.cpp:327-334 (Spiking, STDP), AllDSSynapses.cpp:227-247 (DS,
 * DynamicSTDP). */
static void change_psr(bgo_net *n, size_t i)
{
    const float dt = n->p.dt;
    int k = n->p.syn_kind;
    if (k == BGO_DS || k == BGO_DYNSTDP) {
        if (n->lastSpike[i] != ULONG_MAX64) {
            float isi = (float)(n->step - n->lastSpike[i]) * dt;
            n->r[i] = 1 + (n->r[i] * (1 - n->us[i]) - 1) * n->f_expf(-isi / n->D[i]);
            n->us[i] = n->U[i] + n->us[i] * (1 - n->U[i]) * n->f_expf(-isi / n->F[i]);
        }
        n->psr[i] += ((n->W[i] / n->decay[i]) * n->us[i]) * n->r[i];
        n->lastSpike[i] = n->step;
    } else {
        n->psr[i] += (n->W[i] / n->decay[i]);
    }
}

/* advanceSynapse for one synapse (AllSpikingSynapses.cpp:297-319, AllSTDPSynapses.cpp:283-421);
 * returns the decayed psr the caller adds to the summation point. */
static float advance_synapse(bgo_net *n, size_t i)
{
    const float dt = n->p.dt;
    int k = n->p.syn_kind;
    uint32_t *ring = &n->ring[i * n->W32];
    if (k == BGO_SPIKING || k == BGO_DS) {
        if (ring_test_clear(ring, &n->delayIdx[i], n->R)) change_psr(n, i);
    } else {
        int fPre = ring_test_clear(ring, &n->delayIdx[i], n->R);
        int fPost = ring_test_clear(&n->ringPost[i], &n->delayIdxPost[i], 32);
        if (fPre || fPost) {
            float tauspre = n->tauspre[i], tauspost = n->tauspost[i], taupos = n->taupos[i], tauneg = n->tauneg[i];
            int total_delay = n->total_delay[i];
            int useFD = n->useFD[i];
            int idxPre = n->src[i], idxPost = n->dst[i];
            uint64_t sh, sh2;
            float delta, epre, epost;
            const uint64_t g = n->step;
            const int walk_limit = n->max_spikes;   /* the reference reads outside the ring beyond */
            if (fPre) {
                sh = spike_hist(n, idxPre, -2);
                if (sh != ULONG_MAX64 && useFD) {
                    delta = (float)(uint64_t)(g - sh) * dt;
                    epre = (float)(1.0 - (double)n->f_expf(-delta / tauspre));
                } else {
                    epre = 1.0f;
                }
                for (int off = -1; off >= -walk_limit; --off) {
                    sh = spike_hist(n, idxPost, off);
                    if (sh == ULONG_MAX64) break;
                    /* (spikeHistory - (int64_t)g) is unsigned arithmetic in the reference */
                    delta = (float)(uint64_t)(sh - g) * dt;
                    if ((double)delta <= -3.0 * (double)tauneg) break;
                    if (useFD) {
                        sh2 = spike_hist(n, idxPost, off - 1);
                        if (sh2 == ULONG_MAX64) break;
                        epost = (float)(1.0 - (double)n->f_expf(-((float)(uint64_t)(sh - sh2) * dt) / tauspost));
                    } else {
                        epost = 1.0f;
                    }
                    stdp_learning(n, i, (double)delta, (double)epost, (double)epre);
                }
                change_psr(n, i);
            }
            if (fPost) {
                sh = spike_hist(n, idxPost, -2);
                if (sh != ULONG_MAX64 && useFD) {
                    delta = (float)(uint64_t)(g - sh) * dt;
                    epost = (float)(1.0 - (double)n->f_expf(-delta / tauspost));
                } else {
                    epost = 1.0f;
                }
                for (int off = -1; off >= -walk_limit; --off) {
                    sh = spike_hist(n, idxPre, off);
                    if (sh == ULONG_MAX64) break;
                    delta = (float)(uint64_t)(g - sh - (uint64_t)(int64_t)total_delay) * dt;
                    if (delta <= 0 || (double)delta >= 3.0 * (double)taupos) break;
                    if (useFD) {
                        sh2 = spike_hist(n, idxPre, off - 1);
                        if (sh2 == ULONG_MAX64) break;
                        epre = (float)(1.0 - (double)n->f_expf(-((float)(uint64_t)(sh - sh2) * dt) / tauspre));
                    } else {
                        epre = 1.0f;
                    }
                    stdp_learning(n, i, (double)delta, (double)epost, (double)epre);
                }
            }
        }
    }
    n->psr[i] *= n->decay[i];
    return n->psr[i];
}

/* AllSynapses::advanceSynapses (Synapses/AllSynapses.cpp:325-331) with the summation of
 * AllSpikingSynapses.cpp:314 grouped per destination (the incoming map is destination-major). */
static void advance_synapses(bgo_net *n)
{
    if (!n->has_map) return;
    for (int d = 0; d < n->N; d++) {
        uint32_t b = n->in_begin[d], c = n->in_count[d];
        if (n->p.sum_mode == 0) {
            for (uint32_t q = 0; q < c; q++) n->sp[d] += advance_synapse(n, n->in_map[b + q]);
        } else {
            float acc[64];
            for (int l = 0; l < 64; l++) acc[l] = 0.0f;
            for (uint32_t q = 0; q < c; q++) acc[q & 63] += advance_synapse(n, n->in_map[b + q]);
            for (int off = 32; off >= 1; off >>= 1)
                for (int l = 0; l < off; l++) acc[l] = acc[l] + acc[l + off];
            n->sp[d] += acc[0];
        }
    }
}

/* ------------------------------------------------------------------ stimulus input -------- */
/* SInputPoisson / SInputRegular constructors (Inputs/SInputPoisson.cpp:24-108,
 * Inputs/SInputRegular.cpp:26-140). Called before the model's setup draws. */
static int input_construct(bgo_net *n)
{
    const bgo_params *p = &n->p;
    const int N = n->N;
    const float dt = p->dt;
    if (p->input_kind == BGO_INPUT_POISSON) {
        float fr = p->input_fr_mean / 1000;   /* per ms */
        n->in_lambda = 1 / fr;
        if (p->n_input_mask < 0) {
            for (int i = 0; i < N; i++) n->in_mask[i] = 1;
        } else {
            for (int k = 0; k < p->n_input_mask; k++) {
                int i = p->input_mask[k];
                if (i < 0 || i >= N) return fail(n, "input mask index out of range");
                n->in_mask[i] = 1;
            }
        }
    } else if (p->input_kind == BGO_INPUT_REGULAR) {
        n->in_dur = (int)((double)(p->input_duration / dt) + 0.5);
        int itv = (int)((double)(p->input_interval / dt) + 0.5);
        n->in_cycle = n->in_dur + itv;
        n->in_steps_in_cycle = 0;
        if (n->in_cycle <= 0) return fail(n, "regular input cycle is empty");
        for (int i = 0; i < p->height; i++)
            for (int j = 0; j < p->width; j++)
                n->in_values[i * p->width + j] = p->input_values[(i % 10) * 10 + j % 10];
        if (p->input_sync == BGO_SYNC_NO) {
            for (int i = 0; i < p->height; i++)
                for (int j = 0; j < p->width; j++)
                    n->in_shift[i * p->width + j] = (int)mt_inrange(&n->rng, 0.0f, (float)(n->in_cycle - 1));
        } else if (p->input_sync == BGO_SYNC_WAVE) {
            for (int i = 0; i < p->height; i++)
                for (int j = 0; j < p->width; j++)
                    n->in_shift[i * p->width + j] = (n->in_cycle / p->width) * j;
        }
    } else if (p->input_kind != BGO_INPUT_NONE) {
        return fail(n, "unknown input kind");
    }
    return 0;
}

/* SInputPoisson::init (Inputs/SInputPoisson.cpp:124-146): one AllDSSynapses input synapse per
 * neuron, type EI into inhibitory and EE into excitatory neurons, W = weight * 1e-8. The
 * reference indexes this one-slot-per-neuron object with maxSynapsesPerNeuron * neuron, which is
 * outside its arrays whenever maxSynapsesPerNeuron > 1; slot `neuron` is the intended one. */
static void input_init(bgo_net *n)
{
    if (n->p.input_kind != BGO_INPUT_POISSON) return;
    static const float tau_tab[4] = {(float)6e-3, (float)6e-3, (float)3e-3, (float)3e-3};
    static const float delay_tab[4] = {(float)0.8e-3, (float)0.8e-3, (float)0.8e-3, (float)1.5e-3};
    static const float U_tab[4] = {(float)0.32, (float)0.25, (float)0.05, (float)0.5};
    static const float D_tab[4] = {(float)0.144, (float)0.7, (float)0.125, (float)1.1};
    static const float F_tab[4] = {(float)0.06, (float)0.02, (float)1.2, (float)0.05};
    const float dt = n->p.dt;
    const float unit = (float)1.0e-8;   /* AllSynapses::SYNAPSE_STRENGTH_ADJUSTMENT */
    for (int i = 0; i < n->N; i++) {
        int t = n->ntype[i] == INH ? EI : EE;
        n->in_type[i] = t;
        n->in_W[i] = n->p.input_weight * unit;
        n->in_psr[i] = 0.0f;
        n->in_ring[i] = 0;
        n->in_delayIdx[i] = 0;
        n->in_total_delay[i] = (int)(delay_tab[t] / dt) + 1;
        n->in_decay[i] = expf(-dt / tau_tab[t]);
        n->in_r[i] = 1.0f;
        n->in_u[i] = (float)0.4;
        n->in_lastSpike[i] = ULONG_MAX64;
        n->in_U[i] = U_tab[t]; n->in_D[i] = D_tab[t]; n->in_F[i] = F_tab[t];
    }
}

/* HostSInputPoisson::inputStimulus (Inputs/HostSInputPoisson.cpp:60-94) and
 * HostSInputRegular::inputStimulus (Inputs/HostSInputRegular.cpp:59-80), once per step before the
 * model advance. */
static void input_stimulus(bgo_net *n)
{
    const float dt = n->p.dt;
    if (n->p.input_kind == BGO_INPUT_POISSON) {
        const float lambda = n->in_lambda;
        for (int i = 0; i < n->N; i++) {
            if (!n->in_mask[i]) continue;
            if (--n->in_nisi[i] <= 0) {
                if (ring_insert(&n->in_ring[i], n->in_delayIdx[i], n->in_total_delay[i], 32) < 0) {
                    fail(n, "input delay-queue collision");
                    return;
                }
                if (n->p.noise_mode == 0) {
                    float isi = -lambda * logf(mt_inrange(&n->rng, 0.0f, 1.0f));
                    while (mt_inrange(&n->rng, 0.0f, 1.0f) <= expf(-(isi * isi) / 32))
                        isi = -lambda * logf(mt_inrange(&n->rng, 0.0f, 1.0f));
                    n->in_nisi[i] = (int)((double)((isi / 1000) / dt) + 0.5);
                } else {
                    n->in_nisi[i] = bg_poisson_isi_steps((uint64_t)n->p.seed, (uint32_t)i, n->step, lambda, dt);
                }
            }
            /* AllSpikingSynapses::advanceSynapse with AllDSSynapses::changePSR */
            if (ring_test_clear(&n->in_ring[i], &n->in_delayIdx[i], 32)) {
                if (n->in_lastSpike[i] != ULONG_MAX64) {
                    float isi = (float)(n->step - n->in_lastSpike[i]) * dt;
                    n->in_r[i] = 1 + (n->in_r[i] * (1 - n->in_u[i]) - 1) * n->f_expf(-isi / n->in_D[i]);
                    n->in_u[i] = n->in_U[i] + n->in_u[i] * (1 - n->in_U[i]) * n->f_expf(-isi / n->in_F[i]);
                }
                n->in_psr[i] += ((n->in_W[i] / n->in_decay[i]) * n->in_u[i]) * n->in_r[i];
                n->in_lastSpike[i] = n->step;
            }
            n->in_psr[i] *= n->in_decay[i];
            n->sp[i] += n->in_psr[i];
        }
    } else if (n->p.input_kind == BGO_INPUT_REGULAR) {
        const int c = n->in_steps_in_cycle;
        for (int i = n->N - 1; i >= 0; --i)
            if (c >= n->in_shift[i] && c < (n->in_shift[i] + n->in_dur) % n->in_cycle) n->sp[i] += n->in_values[i];
        n->in_steps_in_cycle = (c + 1) % n->in_cycle;
    }
}

/* ------------------------------------------------------------------ neurons --------------- */
static void fire_record(bgo_net *n, int i)
{   /* AllSpikingNeurons::fire (Neurons/AllSpikingNeurons.cpp:155-173) */
    n->hasFired[i] = 1;
    int idxSp = (n->spikeCount[i] + n->spikeCountOffset[i]) % n->max_spikes;
    n->hist[(size_t)i * n->max_spikes + idxSp] = n->step;
    n->spikeCount[i]++;
}

static inline float draw_noise(bgo_net *n, int i)
{
    if (n->p.noise_mode == 0) return norm_next(&n->norm);
    return bg_noise_normal((uint64_t)n->p.seed, (uint32_t)i, n->step);
}

static void advance_neuron(bgo_net *n, int i)
{
    const float dt = n->p.dt;
    if (n->p.neuron_kind == BGO_LIF) {   /* Neurons/AllLIFNeurons.cpp:20-77 */
        if (n->refr[i] > 0) {
            --n->refr[i];
        } else if (n->Vm[i] >= n->Vthresh[i]) {
            fire_record(n, i);
            n->refr[i] = (int)((double)(n->Trefract[i] / dt) + 0.5);
            n->Vm[i] = n->Vreset[i];
        } else {
            n->sp[i] += n->I0[i];
            float noise = draw_noise(n, i);
            n->sp[i] += noise * n->Inoise[i];
            n->Vm[i] = n->C1[i] * n->Vm[i] + n->C2[i] * n->sp[i];
        }
        n->sp[i] = 0;
    } else {                             /* Neurons/AllIZHNeurons.cpp:570-655 */
        if (n->refr[i] > 0) {
            --n->refr[i];
        } else if (n->Vm[i] >= n->Vthresh[i]) {
            fire_record(n, i);
            n->refr[i] = (int)((double)(n->Trefract[i] / dt) + 0.5);
            n->Vm[i] = (float)((double)n->Cconst[i] * 0.001);
            n->u[i] = n->u[i] + n->Dconst[i];
        } else {
            n->sp[i] += n->I0[i];
            float noise = draw_noise(n, i);
            n->sp[i] += noise * n->Inoise[i];
            float Vint = n->Vm[i] * 1000;
            double t = 0.04 * (double)Vint * (double)Vint + (double)(5 * Vint) + 140 - (double)n->u[i];
            float Vb = (float)((double)Vint + (double)n->C3[i] * t);
            n->u[i] = n->u[i] + n->C3[i] * n->Aconst[i] * (n->Bconst[i] * Vint - n->u[i]);
            n->Vm[i] = (float)((double)Vb * 0.001 + (double)(n->C2[i] * n->sp[i]));
        }
        n->sp[i] = 0;
    }
}

/* AllSpikingNeurons::advanceNeurons (Neurons/AllSpikingNeurons.cpp:98-147) */
static int advance_neurons(bgo_net *n)
{
    int stdp = n->p.syn_kind == BGO_STDP || n->p.syn_kind == BGO_DYNSTDP;
    for (int i = n->N - 1; i >= 0; --i) {
        advance_neuron(n, i);
        if (!n->hasFired[i]) continue;
        if (!(n->spikeCount[i] < n->max_spikes)) return fail(n, "spike history overflow (spikeCount >= max_spikes)");
        if (n->has_map) {
            uint32_t c = n->out_count[i], b = n->out_begin[i];
            for (uint32_t q = 0; q < c; q++) {
                uint32_t s = n->out_map[b + q];
                if (ring_insert(&n->ring[(size_t)s * n->W32], n->delayIdx[s], n->total_delay[s], n->R) < 0)
                    return fail(n, "delay ring collision");
            }
        }
        if (stdp) {
            uint32_t cnt = n->syn_counts[i], notified = 0;
            for (int z = 0; notified < cnt; z++) {
                size_t s = (size_t)n->M * i + z;
                if (n->in_use[s]) {
                    if (ring_insert(&n->ringPost[s], n->delayIdxPost[s], n->total_delayPost[s], 32) < 0)
                        return fail(n, "post delay ring collision");
                    notified++;
                }
            }
        }
        n->hasFired[i] = 0;
    }
    return 0;
}

int bgo_advance(bgo_net *n, int64_t n_steps)
{
    for (int64_t s = 0; s < n_steps; s++) {
        if (n->has_error) return -1;
        input_stimulus(n);                        /* Simulator::advanceUntilGrowth (:177-179) */
        if (advance_neurons(n) < 0) return -1;   /* SingleThreadedSpikingModel::advance (:39-43) */
        advance_synapses(n);
        n->step++;
    }
    return n->has_error ? -1 : 0;
}

/* ------------------------------------------------------------------ growth ---------------- */
/* Overlap area of circles around src (r1) and dest (r2), ConnGrowth::updateOverlap
 * (Connections/ConnGrowth.cpp:337-376) with delta from updateFrontiers (:319-329). */
static float overlap_xy(float xi, float yi, float r1, float xj, float yj, float r2, float (*f_acosf)(float),
                        float (*f_sinf)(float));
static float overlap_area(bgo_net *n, int i, int j)
{
    if (i == j) return 0.0f;   /* delta(i,i) keeps its initial value dist(i,i) = 0 */
    return overlap_xy(n->xloc[i], n->yloc[i], n->radii[i], n->xloc[j], n->yloc[j], n->radii[j], n->f_acosf,
                      n->f_sinf);
}
static float overlap_xy(float xi, float yi, float r1, float xj, float yj, float r2, float (*f_acosf)(float),
                        float (*f_sinf)(float))
{
    float dx = xi - xj, dy = yi - yj;
    float lenAB2 = dx * dx + dy * dy;          /* Layouts/Layout.cpp:59-60 */
    float lenAB = sqrtf(lenAB2);               /* Layouts/Layout.cpp:69 */
    float delta = lenAB - (r1 + r2);
    if (!(delta < 0)) return 0.0f;
    float mn = r1 < r2 ? r1 : r2, mx = r1 < r2 ? r2 : r1;   /* std::min / std::max */
    if (lenAB + mn <= mx) {
        const float pi = (float)3.1415926536;   /* Utils/Global.cpp:90 */
        return pi * mn * mn;
    }
    float r12 = r1 * r1, r22 = r2 * r2;
    float cosCBA = (float)((double)(r22 + lenAB2 - r12) / (2.0 * (double)r2 * (double)lenAB));
    float angCBA = f_acosf(cosCBA);
    float angCBD = (float)(2.0 * (double)angCBA);
    float cosCAB = (float)((double)(r12 + lenAB2 - r22) / (2.0 * (double)r1 * (double)lenAB));
    float angCAB = f_acosf(cosCAB);
    float angCAD = (float)(2.0 * (double)angCAB);
    return (float)(0.5 * (double)(r22 * (angCBD - f_sinf(angCBD)) + r12 * (angCAD - f_sinf(angCAD))));
}

static int apply_weights(bgo_net *n, const float *Wm);

int bgo_update_connections(bgo_net *n)
{
    if (n->has_error) return -1;
    if (n->p.conn_kind != BGO_CONN_GROWTH) return 0;
    const bgo_params *p = &n->p;
    int N = n->N;
    /* ConnGrowth::updateConns (Connections/ConnGrowth.cpp:295-311), VectorMatrix operators in float */
    float maxRate = p->target_rate / p->epsilon;   /* ConnGrowth.cpp:201-203 */
    float c = p->epoch_duration * p->rho;
    for (int i = 0; i < N; i++) {
        if (!(n->spikeCount[i] < n->max_spikes)) return fail(n, "spike count overflow in growth update");
        n->rates[i] = n->spikeCount[i] / p->epoch_duration;
        float t = p->epsilon - n->rates[i] / maxRate;
        t = t / p->beta;
        t = n->f_expf(t);
        t = t + 1.0f;
        t = 2.0f / t;
        float outgrowth = 1.0f - t;
        n->radii[i] += outgrowth * c;
    }
    return apply_weights(n, NULL);
}

/* ConnGrowth::updateSynapsesWeights (Connections/ConnGrowth.cpp:388-463) with W(src, dest) = Wm[src * N
 * + dest], or the circles' overlap area when Wm is NULL (the growth rule's W = area, :395). The
 * reference visits (src, dest) with src outer; every operation touches only dest's slots, so
 * visiting dest outer and src inner performs the same per-destination sequence. An existing
 * synapse with W < 0 is erased (eraseSynapse, Synapses/AllSynapses.cpp:339-344) and its slot is free
 * for a later source's new synapse, as addSynapse takes the first free slot (:357-377). */
static int apply_weights(bgo_net *n, const float *Wm)
{
    int N = n->N, M = n->M;
    const float adj = (float)1.0e-8;   /* AllSynapses::SYNAPSE_STRENGTH_ADJUSTMENT */
    for (int d = 0; d < N; d++) {
        size_t base = (size_t)d * M;
        for (int k = 0; k < M; k++) {
            if (!n->in_use[base + k]) continue;
            int s = n->src[base + k];
            if (n->scratch_slot[s] >= 0) return fail(n, "duplicate (source, destination) synapse");
            n->scratch_slot[s] = k;
        }
        int first_free = 0;
        for (int s = 0; s < N; s++) {
            float area = Wm ? Wm[(size_t)s * N + d] : overlap_area(n, s, d);
            int t = syn_type(n, s, d);
            int k = n->scratch_slot[s];
            if (k >= 0) {
                if (area < 0) {   /* eraseSynapse (AllSynapses.cpp:339-344) */
                    n->syn_counts[d]--;
                    n->in_use[base + k] = 0;
                    n->scratch_slot[s] = -1;
                    if (k < first_free) first_free = k;
                } else {
                    n->W[base + k] = area * syn_sign(t) * adj;
                }
            } else if (area > 0) {
                if ((int)n->syn_counts[d] >= M) return fail(n, "destination neuron has no free synapse slot");
                while (first_free < M && n->in_use[base + first_free]) first_free++;
                int kk = first_free;
                n->syn_counts[d]++;
                if (create_synapse(n, base + kk, s, d, t) < 0) return -1;
                n->W[base + kk] = area * syn_sign(t) * adj;
                n->scratch_slot[s] = kk;
            }
        }
        for (int s = 0; s < N; s++) n->scratch_slot[s] = -1;
    }
    return bgo_rebuild_index_map(n);
}

int bgo_update_weights(bgo_net *n, const float *W)
{
    if (n->has_error) return -1;
    if (!W) return fail(n, "no weight matrix");
    return apply_weights(n, W);
}

/* XmlRecorder::compileHistories (Recorders/XmlRecorder.cpp:71-104) +
 * XmlGrowthRecorder::compileHistories (Recorders/XmlGrowthRecorder.cpp:72-97) +
 * AllSpikingNeurons::clearSpikeCounts (Neurons/AllSpikingNeurons.cpp:79-87). */
int bgo_update_history(bgo_net *n)
{
    if (n->has_error) return -1;
    const float dt = n->p.dt;
    int ms = n->max_spikes;
    for (int i = 0; i < n->N; i++) {
        int cnt = n->spikeCount[i];
        for (int q = 0, idx = n->spikeCountOffset[i]; q < cnt; q++, idx++) {
            if (idx >= ms) idx = 0;
            uint64_t s = n->hist[(size_t)i * ms + idx];
            int idx1 = (int)((double)s * (double)dt);
            int idx2 = (int)((double)s * (double)dt * 100);
            if (idx1 >= 0 && idx1 < n->n_burst) n->burst[idx1] = (float)((double)n->burst[idx1] + 1.0);
            if (idx2 >= 0 && idx2 < n->n_spikesHist) n->spikesHist[idx2] = (float)((double)n->spikesHist[idx2] + 1.0);
        }
    }
    for (int i = 0; i < n->N; i++) {
        n->spikeCountOffset[i] = (n->spikeCount[i] + n->spikeCountOffset[i]) % ms;
        n->spikeCount[i] = 0;
    }
    n->epoch++;
    if (n->p.conn_kind == BGO_CONN_GROWTH && n->epoch <= (n->p.num_epochs > 0 ? n->p.num_epochs : 1)) {
        for (int i = 0; i < n->N; i++) {
            n->ratesHistory[(size_t)n->epoch * n->N + i] = n->rates[i];
            if (n->radii[i] < n->p.min_radius) n->radii[i] = n->p.min_radius;
            n->radiiHistory[(size_t)n->epoch * n->N + i] = n->radii[i];
        }
    }
    return 0;
}

int bgo_run_epoch(bgo_net *n)
{   /* Simulator::advanceUntilGrowth (Core/Simulator.cpp:155-186) then :129-131 */
    int64_t steps = (int64_t)(uint64_t)(n->p.epoch_duration / n->p.dt);
    if (bgo_advance(n, steps) < 0) return -1;
    if (bgo_update_connections(n) < 0) return -1;
    return bgo_update_history(n);
}

void bgo_set_radii(bgo_net *n, float r)
{
    for (int i = 0; i < n->N; i++) n->radii[i] = r;
}

int64_t bgo_step(const bgo_net *n) { return (int64_t)n->step; }
int64_t bgo_total_synapses(const bgo_net *n) { return n->total; }
int32_t bgo_poisson_isi(uint64_t seed, uint32_t neuron, uint64_t step, float lambda_ms, float dt)
{
    return bg_poisson_isi_steps(seed, neuron, step, lambda_ms, dt);
}

const char *bgo_last_error(const bgo_net *n) { return n->has_error ? n->err : ""; }

void *bgo_array(bgo_net *n, const char *name, char *dtype, int64_t *count)
{
    int64_t N = n->N, S = (int64_t)n->N * n->M;
#define A(nm, ptr, dt, cnt) if (!strcmp(name, nm)) { *dtype = dt; *count = (cnt); return (void *)(ptr); }
    A("xloc", n->xloc, 'f', N) A("yloc", n->yloc, 'f', N) A("neuron_type_map", n->ntype, 'i', N)
    A("starter_map", n->starter, 'b', N) A("Vm", n->Vm, 'f', N) A("Vthresh", n->Vthresh, 'f', N)
    A("Vreset", n->Vreset, 'f', N) A("Vrest", n->Vrest, 'f', N) A("Vinit", n->Vinit, 'f', N)
    A("summation_map", n->sp, 'f', N) A("I0", n->I0, 'f', N) A("Inoise", n->Inoise, 'f', N)
    A("Iinject", n->Iinject, 'f', N) A("C1", n->C1, 'f', N) A("C2", n->C2, 'f', N)
    A("Trefract", n->Trefract, 'f', N) A("Aconst", n->Aconst, 'f', N) A("Bconst", n->Bconst, 'f', N)
    A("Cconst", n->Cconst, 'f', N) A("Dconst", n->Dconst, 'f', N) A("u", n->u, 'f', N)
    A("C3", n->C3, 'f', N) A("nStepsInRefr", n->refr, 'i', N) A("hasFired", n->hasFired, 'b', N)
    A("spikeCount", n->spikeCount, 'i', N) A("spikeCountOffset", n->spikeCountOffset, 'i', N)
    A("spike_history", n->hist, 'q', N * n->max_spikes)
    A("in_use", n->in_use, 'b', S) A("sourceNeuronIndex", n->src, 'i', S)
    A("destNeuronIndex", n->dst, 'i', S) A("type", n->type, 'i', S) A("W", n->W, 'f', S)
    A("psr", n->psr, 'f', S) A("decay", n->decay, 'f', S) A("tau", n->tau, 'f', S)
    A("total_delay", n->total_delay, 'i', S) A("delayIdx", n->delayIdx, 'i', S)
    A("delayQueue", n->ring, 'I', S * n->W32) A("delayQueuePost", n->ringPost, 'I', S)
    A("delayIdxPost", n->delayIdxPost, 'i', S) A("r", n->r, 'f', S) A("u_syn", n->us, 'f', S)
    A("U", n->U, 'f', S) A("D", n->D, 'f', S) A("F", n->F, 'f', S) A("lastSpike", n->lastSpike, 'q', S)
    A("Apos", n->Apos, 'f', S) A("Aneg", n->Aneg, 'f', S) A("STDPgap", n->STDPgap, 'f', S)
    A("tauspost", n->tauspost, 'f', S) A("tauspre", n->tauspre, 'f', S) A("taupos", n->taupos, 'f', S)
    A("tauneg", n->tauneg, 'f', S) A("Wex", n->Wex, 'f', S) A("mupos", n->mupos, 'f', S)
    A("muneg", n->muneg, 'f', S) A("useFroemkeDanSTDP", n->useFD, 'b', S)
    A("synapse_counts", n->syn_counts, 'I', N)
    A("incomingSynapseIndexMap", n->in_map, 'I', n->total) A("incomingSynapseBegin", n->in_begin, 'I', N)
    A("incomingSynapseCount", n->in_count, 'I', N) A("outgoingSynapseIndexMap", n->out_map, 'I', n->total)
    A("outgoingSynapseBegin", n->out_begin, 'I', N) A("outgoingSynapseCount", n->out_count, 'I', N)
    A("radii", n->radii, 'f', N) A("rates", n->rates, 'f', N)
    A("radiiHistory", n->radiiHistory, 'f', (int64_t)((n->p.num_epochs > 0 ? n->p.num_epochs : 1) + 1) * N)
    A("ratesHistory", n->ratesHistory, 'f', (int64_t)((n->p.num_epochs > 0 ? n->p.num_epochs : 1) + 1) * N)
    A("in_mask", n->in_mask, 'b', N) A("in_nISIs", n->in_nisi, 'i', N) A("in_type", n->in_type, 'i', N)
    A("in_total_delay", n->in_total_delay, 'i', N) A("in_delayIdx", n->in_delayIdx, 'i', N)
    A("in_shift", n->in_shift, 'i', N) A("in_W", n->in_W, 'f', N) A("in_psr", n->in_psr, 'f', N)
    A("in_decay", n->in_decay, 'f', N) A("in_r", n->in_r, 'f', N) A("in_u", n->in_u, 'f', N)
    A("in_U", n->in_U, 'f', N) A("in_D", n->in_D, 'f', N) A("in_F", n->in_F, 'f', N)
    A("in_values", n->in_values, 'f', N) A("in_lastSpike", n->in_lastSpike, 'q', N)
    A("in_delayQueue", n->in_ring, 'I', N) A("in_steps_in_cycle", &n->in_steps_in_cycle, 'i', 1)
    A("burstinessHist", n->burst, 'f', n->n_burst) A("spikesHistory", n->spikesHist, 'f', n->n_spikesHist)
#undef A
    return NULL;
}

/* ------------------------------------------------------------------ test access ----------- */
float bgo_overlap_area(float x1, float y1, float r1, float x2, float y2, float r2, int math_mode)
{
    return math_mode ? overlap_xy(x1, y1, r1, x2, y2, r2, own_acosf, own_sinf)
                     : overlap_xy(x1, y1, r1, x2, y2, r2, lib_acosf, lib_sinf);
}
void bgo_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { bg_philox4x32_10(ctr, key, out); }
float bgo_noise(uint64_t seed, uint32_t neuron, uint64_t step) { return bg_noise_normal(seed, neuron, step); }
float bgo_norm_stream(uint32_t seed, int n, float *out)
{
    norm_t g;
    norm_seed(&g, seed);
    for (int i = 0; i < n; i++) out[i] = norm_next(&g);
    return n > 0 ? out[n - 1] : 0.0f;
}
double bgo_detmath(int fn, double x, double y)
{
    switch (fn) {
    case 0: return bg_exp(x);
    case 1: return bg_log(x);
    case 2: return bg_sin(x);
    case 3: return bg_cos(x);
    case 4: return bg_acos(x);
    case 5: return bg_pow(x, y);
    case 6: return bg_cos2pi(x);
    default: return (double)bg_expf((float)x);
    }
}
