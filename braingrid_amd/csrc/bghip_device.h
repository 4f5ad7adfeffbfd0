// Device-side network arena and the per-element model arithmetic shared by all kernels.
//
// Layout in HBM (one arena per network):
//   neurons   SoA float/int arrays of length N (Vm, Vthresh, ... exactly the reference's fields)
//   synapses  slot arrays of length S = N*M, destination-major (iSyn = M*dest + k), the layout of
//             Synapses/AllSynapses.cpp:366-371; `meta` packs source | type | total_delay | in_use
//   rings     ring_words x S, word-major (ring[w*S + slot]) so a step touches one coalesced word;
//             for STDP networks byte planes (byte k of a slot at k*S + slot, sring_word), one byte
//   history   uint64 N x max_spikes (Neurons/AllSpikingNeurons.h:185-204)
// Delay rings are kept in absolute-time orientation: the spike delivered at step T sits at bit
// T mod R. The reference's per-synapse delayIdx equals (T - phase) mod R, where phase is the
// creation step mod R; conversion happens at upload/download time only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/bg_detmath.h"

#define BG_META_SRC_MASK 0x3FFFFu
#define BG_META_TYPE_SHIFT 18
#define BG_META_DELAY_SHIFT 20
#define BG_META_IN_USE 0x80000000u
// Implicit-ring networks keep each synapse's cohort (the growth update or synapse-list call that
// created it) in the three meta bits between total_delay and in_use.
#define BG_META_COH_SHIFT 28
#define BG_COH_MAX 8
#define BG_MAX_NEURONS 262144
#define BG_NO_SPIKE 0xFFFFFFFFFFFFFFFFull

enum { NK_LIF = 0, NK_IZH = 1 };
enum { SK_SPIKING = 0, SK_DS = 1, SK_STDP = 2, SK_DYNSTDP = 3 };

// Lists of the spike events of one per-step STDP step, in shards: each shard has its own
// counters, on separate 64-byte lines, and its own segments of the two lists, sized for the worst
// case (every synapse of every destination of the shard), so appends never overflow and concurrent
// waves rarely meet on one counter. A network's runs of BG_STREAM_DESTS destinations go to its
// BG_EV_SHARDS shards in turn. A batch has BG_EV_SHARDS_BATCH shards and gives each member a range
// of its own, wide in proportion to the member's destinations: a shard then lists one network's
// events, so the lanes of an event-pass wavefront share one network descriptor (with more members
// than shards the members' runs go round all shards in turn).
#define BG_EV_SHARDS 128
#define BG_EV_SHARDS_BATCH 1024
#define BG_EV_PAD 16
struct EvLists {
    uint32_t *count;            // destinations of shard s at [s * PAD], synapses at [(SHARDS + s) * PAD]
    const uint32_t *dest_off;   // SHARDS + 1 segment offsets into dest
    const uint32_t *syn_off;    // SHARDS + 1 segment offsets into syn
    uint32_t *dest;             // destinations with events
    uint2 *syn;                 // event synapses as (destination work item, slot)
};

// The STDP parameters createSynapse gives a synapse (Synapses/AllSTDPSynapses.cpp:252-273), kept per
// slot like the reference's arrays and, while every synapse of a type has its type's values, once per
// type: the learning passes then read them from the network descriptor instead of from eleven slot
// arrays (a cache line each per event synapse).
struct StdpParams {
    float Apos, Aneg, gap, tauspost, tauspre, taupos, tauneg, Wex, mupos, muneg;
    int32_t useFD;
};

struct DevNet {
    int N, M, R, W32, max_spikes;
    float dt;
    uint64_t seed;
    uint64_t S;  // N*M
    // neurons
    float *Vm, *Vth, *Vreset, *I0, *Inoise, *C1, *C2, *Trefract, *sp;
    int *refr, *refr_steps;
    float *A, *B, *C, *D, *u, *C3;
    int *cnt, *off;
    uint64_t *hist;
    uint64_t *fired_step;  // last firing step (per-step path), BG_NO_SPIKE if none
    uint32_t *mask[2];     // per-block fire masks (fused path)
    // synapses
    uint32_t *meta;
    uint8_t *phase;
    float *psr, *W;
    uint32_t *ring;
    float *r, *us, *U, *Dd, *F;
    uint64_t *last;
    float *Apos, *Aneg, *gap, *tauspost, *tauspre, *taupos, *tauneg, *Wex, *mupos, *muneg;
    uint8_t *useFD;
    uint8_t *stype;        // STDP: meta's type bits of every slot, for the streaming pass (rebuilt with the
                           //       outgoing map whenever the synapse table changed)
    uint32_t *syn_count, *slot_hi, *in_map, *in_begin;
    // growth
    float *xloc, *yloc, *radii, *rates;
    float *gbox;                 // growth: each 1,024-source tile's min x, max x, min y, max y, max radius
    int32_t *ntype;
    float decay[4];
    int ds_typed;          // 1: every synapse's U, D, F equal its type's creation values
    float tU[4], tD[4], tF[4];
    int stdp_typed;        // 1: every STDP synapse's parameters equal its type's entry of tp
    StdpParams tp[4];
    uint32_t *err;
    // stimulus input (Inputs/SInputPoisson.cpp, Inputs/SInputRegular.cpp)
    int in_kind;                 // BG_INPUT_*
    uint8_t *in_mask;            // Poisson: 0 no input, 1 + k input synapse of kind k (0 EE, 1 EI)
    int32_t *in_nisi;            // Poisson: steps to the next input spike (nISIs)
    float *in_psr, *in_r, *in_u;
    uint64_t *in_last;
    uint32_t *in_ring;           // 32-bit ring per neuron, absolute-time orientation (bit t mod 32)
    float in_W, in_lambda;       // W = weight * 1e-8; mean interval in ms
    float in_decay[2], in_U[2], in_D[2], in_F[2];
    int in_delay[2];
    float *in_values;            // Regular: value per neuron
    int32_t *in_shift;           // Regular: nShiftValues
    int in_dur, in_cycle;        // Regular: nStepsDuration, nStepsCycle
    uint64_t in_cyc_off;         // Regular: nStepsInCycle at step t is (t + in_cyc_off) mod in_cycle
    // per-step STDP path: the step's spike events, in sharded lists (see EvLists)
    EvLists ev;
    const uint32_t *out_begin, *out_count, *out_map;   // outgoing map (synapses of each source)
    uint8_t *out_delay;          // STDP: total delay of each outgoing-map entry's synapse (built with the map)
    const DevNet *self;          // this descriptor in device memory (for kernels that call out of line)
    uint64_t *tbase;             // per-step STDP launches: step = *tbase + the launch's offset (graph replays)
    // Implicit delay rings (Spiking/DS networks with 32-bit rings, fused blocks). A synapse's ring
    // holds exactly the spikes its source fired in the last total_delay steps since the synapse was
    // created, so it is read from the source's recent fire bits instead of being stored; and every
    // synapse of one (source, type, cohort) sees the same deliveries, so their Tsodyks-Markram state
    // (r, u, lastSpike) is one value, advanced by the source's wavefront.
    int impl;                    // 1: this network runs the implicit-ring kernels
    int look;                    // steps of the next block whose first delivery factor is prepared
    uint32_t coh_mask;           // cohort ids in use
    uint64_t young_until;        // blocks starting earlier may see spikes from before a cohort's creation
    int32_t tdelay[4];           // total delay of each synapse type (every synapse has its type's)
    uint32_t *lpt;               // N: destinations by row length, longest first (implicit-ring kernels'
                                 // workgroup order; any order gives the same results)
    uint64_t *rec[2];            // N: fire bits of the 64 steps before a block (bit j = step t0 - 64 + j),
                                 // by block parity like mask[]
    uint64_t *coh_step;          // BG_COH_MAX: creation step of each cohort
    float *cst_r, *cst_u;        // DS: state per (cohort, target class, source) after the last block,
    uint64_t *cst_last;          //     index (c * 2 + class) * N + source; class = type & 1
    float2 *cfac;                // DS: (u, r) after the first delivery of the next block, by block
                                 //     parity: [(parity * BG_COH_MAX + c) * 2 + class] * N + source
    float *Wd;                   // S: W / decay of the synapse's type (the quotient changePSR forms at
                                 //    every delivery), written wherever W is
};

enum { BG_INPUT_NONE = 0, BG_INPUT_POISSON = 1, BG_INPUT_REGULAR = 2 };

__device__ __forceinline__ void dev_error(const DevNet &n, uint32_t bit) { atomicOr(n.err, bit); }

// The delay rings of STDP networks are byte planes: bit b of slot s's ring is bit (b & 7) of byte
// (b >> 3) * S + s of the ring array (the same R / 8 bytes per slot as the 32-bit words w * S + s
// of the Spiking/DS layout). The per-step streaming pass then reads one byte per synapse-step
// instead of a word. A byte shares its 32-bit word with three other slots, so bit updates are
// atomic operations on that word.
__device__ __forceinline__ uint32_t *sring_word(const DevNet &n, uint32_t b, uint64_t slot, uint32_t &bit)
{
    const uint64_t byte = (uint64_t)(b >> 3) * n.S + slot;
    bit = 1u << (((uint32_t)(byte & 3u) << 3) | (b & 7u));
    return (uint32_t *)((uint8_t *)n.ring + (byte & ~3ull));
}

// ---------------------------------------------------------------------------------------------
// Neuron advance. LIF: Neurons/AllLIFNeurons.cpp:20-77; IZH: Neurons/AllIZHNeurons.cpp:570-655;
// spike recording: Neurons/AllSpikingNeurons.cpp:155-173. The summation point value `sp` is the
// synaptic input accumulated in the previous step; the caller zeroes it afterwards.
// Returns true when the neuron fires at step t.
struct NeuronRegs {
    float Vm, Vth, Vreset, I0, Inoise, C1, C2, u, A, B, C, D, C3;
    int refr, refr_steps, cnt, off;
};

template <int NK>
__device__ __forceinline__ bool neuron_advance(const DevNet &n, int i, uint64_t t, float sp, float noise,
                                               NeuronRegs &s)
{
    if (s.refr > 0) {
        --s.refr;
        return false;
    }
    if (s.Vm >= s.Vth) {
        int ms = n.max_spikes;
        int idx = (s.cnt + s.off) % ms;
        n.hist[(uint64_t)i * ms + idx] = t;
        s.cnt++;
        if (!(s.cnt < ms)) dev_error(n, 2u);
        s.refr = s.refr_steps;
        if (NK == NK_LIF) {
            s.Vm = s.Vreset;
        } else {
            s.Vm = (float)((double)s.C * 0.001);
            s.u = s.u + s.D;
        }
        return true;
    }
    float x = sp + s.I0;
    x = x + noise * s.Inoise;
    if (NK == NK_LIF) {
        s.Vm = s.C1 * s.Vm + s.C2 * x;
    } else {
        float Vint = s.Vm * 1000.0f;
        double q = 0.04 * (double)Vint * (double)Vint + (double)(5.0f * Vint) + 140.0 - (double)s.u;
        float Vb = (float)((double)Vint + (double)s.C3 * q);
        s.u = s.u + s.C3 * s.A * (s.B * Vint - s.u);
        s.Vm = (float)((double)Vb * 0.001 + (double)(s.C2 * x));
    }
    return false;
}

// ---------------------------------------------------------------------------------------------
// Stimulus input for neuron i at step t: HostSInputPoisson::inputStimulus
// (Inputs/HostSInputPoisson.cpp:60-94: interval counter, preSpikeHit into the input synapse's
// ring, redraw, then the DS input synapse's advanceSynapse) and HostSInputRegular::inputStimulus
// (Inputs/HostSInputRegular.cpp:59-80: add the value inside the cycle window). Returns whether
// the summation point receives an addition at step t and, in *add, the addend; the caller
// performs the reference's single `summationPoint += add`. The input does not depend on the
// neuron's own state, so the fused kernels evaluate a whole block of steps ahead of the neuron
// loop.
struct InputRegs {
    uint32_t m;        // Poisson: mask byte
    int nisi;
    uint32_t ring;
    float psr, r, u;
    uint64_t last;
    float value;       // Regular
    int shift;
};

__device__ __forceinline__ void input_load(const DevNet &n, int i, InputRegs &q)
{
    q.m = 0;
    if (n.in_kind == BG_INPUT_POISSON) {
        q.m = n.in_mask[i];
        if (q.m) {
            q.nisi = n.in_nisi[i]; q.ring = n.in_ring[i]; q.psr = n.in_psr[i];
            q.r = n.in_r[i]; q.u = n.in_u[i]; q.last = n.in_last[i];
        }
    } else if (n.in_kind == BG_INPUT_REGULAR) {
        q.value = n.in_values[i];
        q.shift = n.in_shift[i];
    }
}

__device__ __forceinline__ void input_store(const DevNet &n, int i, const InputRegs &q)
{
    if (n.in_kind == BG_INPUT_POISSON && q.m) {
        n.in_nisi[i] = q.nisi; n.in_ring[i] = q.ring; n.in_psr[i] = q.psr;
        n.in_r[i] = q.r; n.in_u[i] = q.u; n.in_last[i] = q.last;
    }
}

__device__ __forceinline__ bool input_step_regs(const DevNet &n, int i, uint64_t t, InputRegs &q, float *add)
{
    if (n.in_kind == BG_INPUT_POISSON) {
        if (q.m == 0) return false;
        const int k = (int)q.m - 1;
        const uint32_t tb = (uint32_t)t & 31u;
        if (--q.nisi <= 0) {
            const uint32_t b = 1u << ((tb + (uint32_t)n.in_delay[k]) & 31u);
            if (q.ring & b) dev_error(n, 1u);
            q.ring |= b;
            q.nisi = bg_poisson_isi_steps(n.seed, (uint32_t)i, t, n.in_lambda, n.dt);
        }
        const float decay = n.in_decay[k];
        if ((q.ring >> tb) & 1u) {
            q.ring &= ~(1u << tb);
            if (q.last != BG_NO_SPIKE) {
                const float isi = (float)(t - q.last) * n.dt;
                q.r = 1.0f + (q.r * (1.0f - q.u) - 1.0f) * bg_expf(-isi / n.in_D[k]);
                q.u = n.in_U[k] + q.u * (1.0f - n.in_U[k]) * bg_expf(-isi / n.in_F[k]);
            }
            q.last = t;
            q.psr = q.psr + ((n.in_W / decay) * q.u) * q.r;
        }
        q.psr = q.psr * decay;
        *add = q.psr;
        return true;
    }
    if (n.in_kind == BG_INPUT_REGULAR) {
        const int c = (int)((t + n.in_cyc_off) % (uint64_t)n.in_cycle);
        if (c >= q.shift && c < (q.shift + n.in_dur) % n.in_cycle) {
            *add = q.value;
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ bool input_step(const DevNet &n, int i, uint64_t t, float *add)
{
    InputRegs q;
    input_load(n, i, q);
    const bool r = input_step_regs(n, i, t, q, add);
    input_store(n, i, q);
    return r;
}

template <int NK>
__device__ __forceinline__ void neuron_load(const DevNet &n, int i, NeuronRegs &s)
{
    s.Vm = n.Vm[i]; s.Vth = n.Vth[i]; s.Vreset = n.Vreset[i]; s.I0 = n.I0[i]; s.Inoise = n.Inoise[i];
    s.C1 = n.C1[i]; s.C2 = n.C2[i]; s.refr = n.refr[i]; s.refr_steps = n.refr_steps[i];
    s.cnt = n.cnt[i]; s.off = n.off[i];
    if (NK == NK_IZH) { s.u = n.u[i]; s.A = n.A[i]; s.B = n.B[i]; s.C = n.C[i]; s.D = n.D[i]; s.C3 = n.C3[i]; }
}

template <int NK>
__device__ __forceinline__ void neuron_store(const DevNet &n, int i, const NeuronRegs &s)
{
    n.Vm[i] = s.Vm; n.refr[i] = s.refr; n.cnt[i] = s.cnt;
    if (NK == NK_IZH) n.u[i] = s.u;
}

// ---------------------------------------------------------------------------------------------
// The amount a delivered spike adds to psr, with the synapse-state update that goes with it:
// AllDSSynapses::changePSR (Synapses/AllDSSynapses.cpp:227-247) for DS / DynamicSTDP,
// AllSpikingSynapses::changePSR (:327-334) otherwise. The caller adds the result to psr, which
// is the same single float addition the reference's `psr += ...` performs.
template <int SK>
__device__ __forceinline__ float psr_increment(const DevNet &n, uint64_t slot, int type, uint64_t t, float decay)
{
    if (SK == SK_DS || SK == SK_DYNSTDP) {
        float r = n.r[slot], u = n.us[slot];
        uint64_t last = n.last[slot];
        if (last != BG_NO_SPIKE) {
            float isi = (float)(t - last) * n.dt;
            float D = n.ds_typed ? n.tD[type] : n.Dd[slot];
            float F = n.ds_typed ? n.tF[type] : n.F[slot];
            float U = n.ds_typed ? n.tU[type] : n.U[slot];
            r = 1.0f + (r * (1.0f - u) - 1.0f) * bg_expf(-isi / D);
            u = U + u * (1.0f - U) * bg_expf(-isi / F);
            n.r[slot] = r;
            n.us[slot] = u;
        }
        n.last[slot] = t;
        return ((n.W[slot] / decay) * u) * r;
    } else {
        return n.W[slot] / decay;
    }
}

template <int SK>
__device__ __forceinline__ float change_psr(const DevNet &n, uint64_t slot, int type, uint64_t t, float psr,
                                            float decay)
{
    return psr + psr_increment<SK>(n, slot, type, t, decay);
}

// getSpikeHistory (Neurons/AllSpikingNeurons.cpp:182-188)
__device__ __forceinline__ uint64_t spike_hist(const DevNet &n, int idx, int offIndex)
{
    int ms = n.max_spikes;
    int k = (n.cnt[idx] + n.off[idx] + ms + offIndex) % ms;
    return n.hist[(uint64_t)idx * ms + k];
}

// A synapse's STDP parameters: its type's when the network is typed, else its slot's.
__device__ __forceinline__ StdpParams stdp_params(const DevNet &n, uint64_t slot, int type)
{
    if (n.stdp_typed) return n.tp[type];
    StdpParams p;
    p.Apos = n.Apos[slot]; p.Aneg = n.Aneg[slot]; p.gap = n.gap[slot]; p.tauspost = n.tauspost[slot];
    p.tauspre = n.tauspre[slot]; p.taupos = n.taupos[slot]; p.tauneg = n.tauneg[slot]; p.Wex = n.Wex[slot];
    p.mupos = n.mupos[slot]; p.muneg = n.muneg[slot]; p.useFD = n.useFD[slot];
    return p;
}

// AllSTDPSynapses::stdpLearning (Synapses/AllSTDPSynapses.cpp:432-472)
__device__ __forceinline__ void stdp_learning(const DevNet &n, uint64_t slot, const StdpParams &P, double delta,
                                              double epost, double epre)
{
    const float STDPgap = P.gap, muneg = P.muneg, mupos = P.mupos, tauneg = P.tauneg;
    const float taupos = P.taupos, Aneg = P.Aneg, Apos = P.Apos, Wex = P.Wex;
    float W = n.W[slot];
    float dw;
    if (delta < -(double)STDPgap) {
        dw = (float)((double)(bg_powf(W, muneg) * Aneg) * bg_exp(delta / (double)tauneg));
    } else if (delta > (double)STDPgap) {
        dw = (float)((double)(bg_powf(Wex - W, mupos) * Apos) * bg_exp(-delta / (double)taupos));
    } else {
        return;
    }
    W = (float)((double)W + epost * epre * (double)dw);
    if ((Wex < 0 && W > 0) || (Wex > 0 && W < 0)) W = 0;
    if (fabsf(W) > fabsf(Wex)) W = Wex;
    n.W[slot] = W;
}

// The learning part of AllSTDPSynapses::advanceSynapse (Synapses/AllSTDPSynapses.cpp:289-407).
// The pre-spike walk over the post neuron's history computes (spikeHistory - (int64_t)step) in
// unsigned 64-bit arithmetic, exactly as the reference does, so that walk never meets its
// `delta <= -3 tauneg` exit and runs over the whole history ring; both walks are bounded by the
// ring length max_spikes, beyond which the reference would index outside the history ring.
template <int SK>
__device__ __noinline__ float stdp_event(const DevNet &n, uint64_t slot, int type, uint64_t t, bool fPre, bool fPost,
                                         int idxPre, int idxPost, int total_delay, float psr, float decay)
{
    const float dt = n.dt;
    const StdpParams P = stdp_params(n, slot, type);
    const float tauspre = P.tauspre, tauspost = P.tauspost, taupos = P.taupos, tauneg = P.tauneg;
    const bool useFD = P.useFD != 0;
    const int walk = n.max_spikes;
    uint64_t sh, sh2;
    float delta, epre, epost;
    if (fPre) {
        sh = spike_hist(n, idxPre, -2);
        if (sh != BG_NO_SPIKE && useFD) {
            delta = (float)(uint64_t)(t - sh) * dt;
            epre = (float)(1.0 - (double)bg_expf(-delta / tauspre));
        } else {
            epre = 1.0f;
        }
        for (int o = -1; o >= -walk; --o) {
            sh = spike_hist(n, idxPost, o);
            if (sh == BG_NO_SPIKE) break;
            delta = (float)(uint64_t)(sh - t) * dt;
            if ((double)delta <= -3.0 * (double)tauneg) break;
            if (useFD) {
                sh2 = spike_hist(n, idxPost, o - 1);
                if (sh2 == BG_NO_SPIKE) break;
                epost = (float)(1.0 - (double)bg_expf(-((float)(uint64_t)(sh - sh2) * dt) / tauspost));
            } else {
                epost = 1.0f;
            }
            const uint32_t w0 = __float_as_uint(n.W[slot]);
            stdp_learning(n, slot, P, (double)delta, (double)epost, (double)epre);
            // An entry older than step t has an interval that wraps in the unsigned arithmetic to
            // (float)2^64 * dt, and so does every older entry after it (all are earlier steps, and
            // t < 2^39 keeps the float conversion at 2^64). With that delta, dw is P(W) * 0, so a
            // pass that left W bit-identical (and not NaN) leaves it so for every older entry:
            // the rest of the walk cannot change anything and is skipped.
            if (sh < t && t < (1ull << 39)) {
                const float w1 = n.W[slot];
                if (__float_as_uint(w1) == w0 && w1 == w1) break;
            }
        }
        psr = change_psr<SK == SK_DYNSTDP ? SK_DS : SK_SPIKING>(n, slot, type, t, psr, decay);
    }
    if (fPost) {
        sh = spike_hist(n, idxPost, -2);
        if (sh != BG_NO_SPIKE && useFD) {
            delta = (float)(uint64_t)(t - sh) * dt;
            epost = (float)(1.0 - (double)bg_expf(-delta / tauspost));
        } else {
            epost = 1.0f;
        }
        for (int o = -1; o >= -walk; --o) {
            sh = spike_hist(n, idxPre, o);
            if (sh == BG_NO_SPIKE) break;
            delta = (float)(uint64_t)(t - sh - (uint64_t)(int64_t)total_delay) * dt;
            if (delta <= 0 || (double)delta >= 3.0 * (double)taupos) break;
            if (useFD) {
                sh2 = spike_hist(n, idxPre, o - 1);
                if (sh2 == BG_NO_SPIKE) break;
                epre = (float)(1.0 - (double)bg_expf(-((float)(uint64_t)(sh - sh2) * dt) / tauspre));
            } else {
                epre = 1.0f;
            }
            stdp_learning(n, slot, P, (double)delta, (double)epost, (double)epre);
        }
    }
    return psr;
}

// One event synapse of the per-step STDP path, whole: the ring test and clear, the learning walks
// of stdp_event, changePSR (AllDSSynapses::changePSR for DynamicSTDP, W / decay for STDP) and the
// decay and store of the PSR (Synapses/AllSTDPSynapses.cpp:289-421), with stdp_event's and
// stdp_learning's arithmetic, in the same order. The event pass is bound by the latency of one such
// chain per thread, so the body requests everything it reads before its first store: the slot's
// state and parameters, the ring word, both neurons' history cursors and then the four newest
// entries of both histories (the walks rarely go further; deeper entries are read when reached). W,
// r, u and lastSpike are carried in registers through the walks and stored once at the end.
// stdp_event reads each field at first use and re-reads W after every store, a chain several
// times longer.
template <int SK>
__device__ __noinline__ void stdp_event_slot(const DevNet &n, int d, uint64_t slot, uint64_t t)
{
    // ---- reads: slot and destination first, then what depends on meta, then history entries
    const int ms = n.max_spikes;
    const float dt = n.dt;
    const uint32_t tb = (uint32_t)(t & (uint64_t)(n.R - 1));
    uint32_t m;
    uint32_t *const wp = sring_word(n, tb, slot, m);
    float *const Wp = &n.W[slot];
    float *const psrp = &n.psr[slot];
    const uint64_t *const hist = n.hist;
    const uint32_t word = *wp;
    const uint64_t fs_post = n.fired_step[d];
    const uint32_t meta = n.meta[slot];
    float psr = *psrp;
    float W = *Wp;
    const int base_post = n.cnt[d] + n.off[d] + ms;
    float r = 0.0f, u = 0.0f, Us = 0.0f, Ds = 0.0f, Fs = 0.0f;
    uint64_t last = BG_NO_SPIKE;
    const bool typed = n.ds_typed != 0;
    float *const rp = SK == SK_DYNSTDP ? &n.r[slot] : nullptr;
    float *const up = SK == SK_DYNSTDP ? &n.us[slot] : nullptr;
    uint64_t *const lastp = SK == SK_DYNSTDP ? &n.last[slot] : nullptr;
    if (SK == SK_DYNSTDP) {
        r = *rp;
        u = *up;
        last = *lastp;
        if (!typed) {
            Ds = n.Dd[slot];
            Fs = n.F[slot];
            Us = n.U[slot];
        }
    }
    const int src = (int)(meta & BG_META_SRC_MASK);
    const int type = (meta >> BG_META_TYPE_SHIFT) & 3;
    const int total_delay = (int)((meta >> BG_META_DELAY_SHIFT) & 0xFFu);
    const float decay = n.decay[type];
    const StdpParams P = stdp_params(n, slot, type);
    const float tauspre = P.tauspre, tauspost = P.tauspost, taupos = P.taupos, tauneg = P.tauneg;
    const bool useFD = P.useFD != 0;
    const float gap = P.gap, muneg = P.muneg, mupos = P.mupos, Aneg = P.Aneg, Apos = P.Apos, Wex = P.Wex;
    if (SK == SK_DYNSTDP && typed) {
        Ds = n.tD[type];
        Fs = n.tF[type];
        Us = n.tU[type];
    }
    const int base_pre = n.cnt[src] + n.off[src] + ms;
    // entries o = -1 .. -4 of both histories; an entry is at (cnt + off + ms + o) mod ms, and
    // cnt + off + ms - 1 >= 0 (count and offset are not negative)
    const int kpost = (base_post - 1) % ms, kpre = (base_pre - 1) % ms;
    auto at = [&](int k1, int j) {
        int k = k1 - j;
        while (k < 0) k += ms;
        return k;
    };
    const uint64_t *const hpo = hist + (uint64_t)d * ms, *const hpr = hist + (uint64_t)src * ms;
    const uint64_t po1 = hpo[kpost], po2 = hpo[at(kpost, 1)], po3 = hpo[at(kpost, 2)], po4 = hpo[at(kpost, 3)];
    const uint64_t pr1 = hpr[kpre], pr2 = hpr[at(kpre, 1)], pr3 = hpr[at(kpre, 2)], pr4 = hpr[at(kpre, 3)];
    // getSpikeHistory(idx, o): a prefetched entry where the reference's index agrees with it, else
    // the reference's own index expression
    auto hget = [&](bool post_side, int o) -> uint64_t {
        const int base = post_side ? base_post : base_pre;
        if (o >= -4 && base + o >= 0) {
            const uint64_t a = post_side ? po1 : pr1, b = post_side ? po2 : pr2;
            const uint64_t c = post_side ? po3 : pr3, e = post_side ? po4 : pr4;
            return o == -1 ? a : (o == -2 ? b : (o == -3 ? c : e));
        }
        const int idx = post_side ? d : src;
        return hist[(uint64_t)idx * ms + (base + o) % ms];
    };
    // AllSTDPSynapses::stdpLearning on the W in registers
    bool learned = false;
    auto learn = [&](double delta, double epost, double epre) {
        float dw;
        if (delta < -(double)gap) {
            dw = (float)((double)(bg_powf(W, muneg) * Aneg) * bg_exp(delta / (double)tauneg));
        } else if (delta > (double)gap) {
            dw = (float)((double)(bg_powf(Wex - W, mupos) * Apos) * bg_exp(-delta / (double)taupos));
        } else {
            return;
        }
        float Wn = (float)((double)W + epost * epre * (double)dw);
        if ((Wex < 0 && Wn > 0) || (Wex > 0 && Wn < 0)) Wn = 0;
        if (fabsf(Wn) > fabsf(Wex)) Wn = Wex;
        W = Wn;
        learned = true;
    };

    // ---- the event
    const bool pre = (word & m) != 0;
    const bool post = fs_post == t;
    bool ru_changed = false;
    uint64_t sh;
    float delta, epre, epost;
    if (pre) {
        sh = hget(false, -2);
        if (sh != BG_NO_SPIKE && useFD) {
            delta = (float)(uint64_t)(t - sh) * dt;
            epre = (float)(1.0 - (double)bg_expf(-delta / tauspre));
        } else {
            epre = 1.0f;
        }
        for (int o = -1; o >= -ms; --o) {
            sh = hget(true, o);
            if (sh == BG_NO_SPIKE) break;
            delta = (float)(uint64_t)(sh - t) * dt;
            if ((double)delta <= -3.0 * (double)tauneg) break;
            if (useFD) {
                const uint64_t sh2 = hget(true, o - 1);
                if (sh2 == BG_NO_SPIKE) break;
                epost = (float)(1.0 - (double)bg_expf(-((float)(uint64_t)(sh - sh2) * dt) / tauspost));
            } else {
                epost = 1.0f;
            }
            const uint32_t w0 = __float_as_uint(W);
            learn((double)delta, (double)epost, (double)epre);
            // an older entry that left W bit-identical: so do all older ones (see stdp_event)
            if (sh < t && t < (1ull << 39) && __float_as_uint(W) == w0 && W == W) break;
        }
        if (SK == SK_DYNSTDP) {
            if (last != BG_NO_SPIKE) {
                const float isi = (float)(t - last) * dt;
                r = 1.0f + (r * (1.0f - u) - 1.0f) * bg_expf(-isi / Ds);
                u = Us + u * (1.0f - Us) * bg_expf(-isi / Fs);
                ru_changed = true;
            }
            last = t;
            psr = psr + ((W / decay) * u) * r;
        } else {
            psr = psr + W / decay;
        }
    }
    if (post) {
        sh = hget(true, -2);
        if (sh != BG_NO_SPIKE && useFD) {
            delta = (float)(uint64_t)(t - sh) * dt;
            epost = (float)(1.0 - (double)bg_expf(-delta / tauspost));
        } else {
            epost = 1.0f;
        }
        for (int o = -1; o >= -ms; --o) {
            sh = hget(false, o);
            if (sh == BG_NO_SPIKE) break;
            delta = (float)(uint64_t)(t - sh - (uint64_t)(int64_t)total_delay) * dt;
            if (delta <= 0 || (double)delta >= 3.0 * (double)taupos) break;
            if (useFD) {
                const uint64_t sh2 = hget(false, o - 1);
                if (sh2 == BG_NO_SPIKE) break;
                epre = (float)(1.0 - (double)bg_expf(-((float)(uint64_t)(sh - sh2) * dt) / tauspre));
            } else {
                epre = 1.0f;
            }
            learn((double)delta, (double)epost, (double)epre);
        }
    }

    // ---- stores
    if (pre) atomicAnd(wp, ~m);
    if (learned) *Wp = W;
    if (SK == SK_DYNSTDP && pre) {
        if (ru_changed) {
            *rp = r;
            *up = u;
        }
        *lastp = t;
    }
    *psrp = psr * decay;
}

// Wavefront sum in the engine's canonical order: lane l holds the lane-strided sequential partial
// sum; butterfly over 32, 16, ..., 1 leaves the same value in every lane (for lane 0 this is the
// tree acc[l] += acc[l + off] of the oracle's wave64 mode).
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

// ConnGrowth::updateFrontiers + updateOverlap for one (i = src, j = dest) pair
// (Connections/ConnGrowth.cpp:319-376, Layouts/Layout.cpp:54-69).
__device__ __forceinline__ float overlap_area(float xi, float yi, float r1, float xj, float yj, float r2, bool same)
{
    if (same) return 0.0f;
    float dx = xi - xj, dy = yi - yj;
    float lenAB2 = dx * dx + dy * dy;
    float lenAB = sqrtf(lenAB2);
    float delta = lenAB - (r1 + r2);
    if (!(delta < 0)) return 0.0f;
    float mn = r1 < r2 ? r1 : r2, mx = r1 < r2 ? r2 : r1;
    if (lenAB + mn <= mx) {
        const float pi = 3.1415926536f;
        return pi * mn * mn;
    }
    float r12 = r1 * r1, r22 = r2 * r2;
    float cosCBA = (float)((double)(r22 + lenAB2 - r12) / (2.0 * (double)r2 * (double)lenAB));
    float angCBA = bg_acosf(cosCBA);
    float angCBD = (float)(2.0 * (double)angCBA);
    float cosCAB = (float)((double)(r12 + lenAB2 - r22) / (2.0 * (double)r1 * (double)lenAB));
    float angCAB = bg_acosf(cosCAB);
    float angCAD = (float)(2.0 * (double)angCAB);
    return (float)(0.5 * (double)(r22 * (angCBD - bg_sinf(angCBD)) + r12 * (angCAD - bg_sinf(angCAD))));
}
