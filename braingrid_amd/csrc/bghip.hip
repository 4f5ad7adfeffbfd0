// libbghip.so: the C ABI of include/bghip.h on top of the kernels in bghip_kernels.h.
//
// One network = one HIP stream + one device arena (a single hipMalloc carved into the SoA
// arrays). The host side converts between the reference's array conventions (relative delay-ring
// orientation, per-synapse delayIdx, separate type/source/delay arrays) and the device's packed
// slot layout only at upload and download time.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <new>
#include <vector>
#include "../../include/bghip.h"
#include "bghip_kernels_epoch.h"
#include "bghip_implicit.h"
#include "bghip_launch.h"

// Scratch memory of bghip_update_weights_static: the caller's N x N weight matrix for N up to
// BG_SCRATCH_MAX_N. One array per GPU, held until the process ends. In host code `scratch` only
// names the symbol; its device address on each GPU is looked up once per device
// (hipGetSymbolAddress, at context creation with that device current) and only that address is
// passed to copies and kernels.
#define BG_SCRATCH_MAX_N 2048
#define BG_SCRATCH_BYTES ((size_t)BG_SCRATCH_MAX_N * BG_SCRATCH_MAX_N * sizeof(float))
#define BG_MAX_DEVICES 64
static __device__ __align__(256) char scratch[BG_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[BG_MAX_DEVICES];
static hipError_t g_scratch_err[BG_MAX_DEVICES];

// The device address of `scratch` on the current device `device`, looked up on the first call.
static void *scratch_address(int device)
{
    if (device < 0 || device >= BG_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device] && g_scratch_err[device] == hipSuccess) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) g_scratch_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        else g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// A launch shape the batch autotuner picked, kept per device for groups of the same kind and
// size class (work items within a factor of two), so re-formed groups do not repeat the trials.
struct TuneEntry {
    int mode, nk, sk, w32, size_class;
    BatchShape shape;
    float step_us;
};

struct bghip_ctx {
    int device;
    char err[512];
    std::vector<TuneEntry> tuned;
    // per-step runs replay captured graphs; BGHIP_GRAPHS=0 in the environment launches the same
    // kernels directly instead (for tools that do not follow graph launches)
    bool graphs;
    // batches of this context: a destroyed network leaves every batch it is a member of
    std::vector<bghip_batch *> batches;
    // Spiking/DS networks with 32-bit rings run the implicit-ring kernels; BGHIP_IMPLICIT=0 in the
    // environment keeps them on the explicit ones (for comparisons)
    bool implicit;
    void *scratch;   // device address of the static scratch array on this device (null: unavailable)
};

struct bghip_net {
    bghip_ctx *ctx;
    bghip_net_desc desc;
    DevNet dn;
    hipStream_t stream;
    char *arena;
    size_t arena_bytes;
    uint64_t step;
    int kmax;              // steps per fused launch
    int sel;               // mask buffer holding the most recent block's fire masks
    uint64_t last_t0;
    bool pending;          // most recent block's spikes not yet inserted into the rings
    int min_delay_uploaded;
    bool has_layout, has_growth;
    bool tp_set[4];                // STDP: the entry of each type in dn.tp has been set
    std::vector<int32_t> ntype_host;   // the uploaded layout's neuron types (synapse types of added synapses)
    bghip_growth growth;
    uint32_t *d_out_map, *d_out_begin, *d_out_count, *d_cursor, *d_stats;
    uint64_t *d_total;
    uint64_t total;
    bool out_valid;
    hipEvent_t ev0, ev1;
    bool timing;
    int64_t timed_launches;
    hipEvent_t ext_ev;     // work another stream (a batch) must finish before this net's stream runs
    bool ext_pending;
    uint32_t in_phase;     // Poisson input ring: reference delayIdx = (step - in_phase) mod 32
    int32_t *in_idx_host;  // N: delayIdx of neurons without input (their synapse never advances)
    uint8_t *in_mask_host; // N: copy of dn.in_mask
    uint64_t *h_total;     // pinned landing slot of the synapse total after an index-map rebuild
    uint32_t *h_err;       // pinned landing slot of the device error bits (read by bghip_sync)
    bool total_pending;    // h_total not read yet (the rebuild is still in flight)
    // on-device recorder (XmlRecorder / XmlGrowthRecorder histories)
    uint32_t *d_burst, *d_shist;    // spike counts per 1 s and 10 ms bin
    float *d_rates_hist, *d_radii_hist;
    int64_t n_burst, n_shist;
    int32_t rec_epochs, rec_epoch;  // rows (num_epochs + 1) and rows written so far
    bool rec_on, rec_growth;
    uint64_t hist_step0;            // first step whose spikes the next compile bins
    bool hist_range_known;
    DevNet *d_self;        // device copy of dn (dn.self), refreshed when dn changes
    uint32_t *d_ev_off;    // per-step STDP: shard offsets of the event lists (dest, then syn)
    DevNet dn_shadow;      // dn as last copied to d_self
    // per-step STDP: BG_GRAPH_STEPS steps captured as one graph, valid while dn equals g_dn
    hipGraphExec_t g_exec = nullptr;
    DevNet g_dn;
    // implicit rings: cohort ids handed out (coh_alloc), those handed out since the last usage scan
    // was issued (coh_since), and that scan's result landing in pinned memory
    uint32_t coh_alloc, coh_since;
    bool coh_scan_pending;
    uint32_t *h_coh_used, *d_coh_used;
    hipEvent_t coh_ev;
};

static int to_explicit(bghip_net *n);
static uint64_t net_total(bghip_net *n);

// Orders the network's own stream after outstanding batch work on it.
static void net_begin(bghip_net *n)
{
    if (n->ext_pending) {
        (void)hipStreamWaitEvent(n->stream, n->ext_ev, 0);
        n->ext_pending = false;
    }
}

static int set_err(bghip_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(ctx->err, sizeof ctx->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIP_TRY(ctx, call)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_err((ctx), BGHIP_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------------------------------------
extern "C" int bghip_ctx_create(int device, bghip_ctx **out)
{
    if (!out) return BGHIP_EINVAL;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return BGHIP_EDEVICE;
    if (hipSetDevice(device) != hipSuccess) return BGHIP_EDEVICE;
    bghip_ctx *c = new (std::nothrow) bghip_ctx();
    if (!c) return BGHIP_ENOMEM;
    c->device = device;
    c->err[0] = 0;
    const char *g = getenv("BGHIP_GRAPHS");
    c->graphs = !(g && strcmp(g, "0") == 0);
    const char *im = getenv("BGHIP_IMPLICIT");
    c->implicit = !(im && strcmp(im, "0") == 0);
    c->scratch = scratch_address(device);
    *out = c;
    return BGHIP_OK;
}

extern "C" void bghip_ctx_destroy(bghip_ctx *ctx) { delete ctx; }
extern "C" const char *bghip_last_error(const bghip_ctx *ctx) { return ctx ? ctx->err : "no context"; }

extern "C" int bghip_ctx_mem_info(bghip_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes)
{
    size_t f = 0, t = 0;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return BGHIP_OK;
}

// ------------------------------------------------------------------------------------------------
// Arena layout. Every array is 256-byte aligned.
struct ArenaPlan {
    size_t off = 0;
    size_t take(size_t bytes)
    {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    }
};

static bool desc_ok(const bghip_net_desc *d)
{
    if (!d || d->num_neurons <= 0 || d->num_neurons > BG_MAX_NEURONS) return false;
    if (d->max_synapses_per_neuron < 0 || d->max_spikes <= 0) return false;
    int R = d->ring_bits;
    if (R != 32 && R != 64 && R != 128 && R != 256) return false;
    if (d->neuron_model < 0 || d->neuron_model > 1 || d->synapse_model < 0 || d->synapse_model > 3) return false;
    if (!(d->dt > 0)) return false;
    for (int t = 0; t < 4; t++)
        if (d->types.total_delay[t] < 1 || d->types.total_delay[t] >= R || d->types.total_delay[t] > 255) return false;
    return true;
}

struct Carve {
    ArenaPlan plan;
    std::vector<std::pair<void **, size_t>> items;
    template <typename T>
    void add(T **p, size_t count)
    {
        items.push_back({(void **)p, plan.take(sizeof(T) * (count ? count : 1))});
    }
};

// Segment sizes of the sharded event lists (NS shards). The streaming pass works on runs ("chunks")
// of BG_STREAM_DESTS destinations; a part is one network, whose chunk j lists its events in shard
// (sbase + j mod m) mod NS. Each shard gets room for the worst case: every destination of its
// chunks listed, with all M of its slots. Returns false when the synapse list would not be
// addressable with 32-bit offsets.
struct EvPart { uint64_t N, M, sbase, m; };
static uint32_t ev_shard(const EvPart &p, uint64_t chunk, uint64_t NS) { return (uint32_t)((p.sbase + chunk % p.m) % NS); }
static bool ev_offsets(const std::vector<EvPart> &parts, uint64_t NS, std::vector<uint32_t> &dest_off,
                       std::vector<uint32_t> &syn_off)
{
    const uint64_t D = BG_STREAM_DESTS;
    std::vector<uint64_t> dc(NS, 0), sc(NS, 0);
    for (const EvPart &p : parts)
        for (uint64_t k = 0; k * D < p.N; k++) {
            const uint64_t r = ev_shard(p, k, NS), nd = std::min<uint64_t>(D, p.N - k * D);
            dc[r] += nd;
            sc[r] += nd * p.M;
        }
    dest_off.assign(NS + 1, 0);
    syn_off.assign(NS + 1, 0);
    uint64_t ad = 0, as = 0;
    for (uint64_t r = 0; r < NS; r++) {
        dest_off[r] = (uint32_t)ad;
        syn_off[r] = (uint32_t)as;
        ad += dc[r];
        as += sc[r];
    }
    dest_off[NS] = (uint32_t)ad;
    syn_off[NS] = (uint32_t)as;
    return as < (1ull << 32);
}

// The batch's shard ranges: member k gets m_k = 1 + (NS - n) * chunks_k / chunks shards from
// sbase_k on (so sum m_k <= NS, and a larger member spreads its list appends over more counters);
// with more members than shards every member's chunks go round all NS shards, continuing where the
// previous member's ended.
static void ev_batch_parts(std::vector<EvPart> &parts, uint64_t NS)
{
    const uint64_t D = BG_STREAM_DESTS, n = parts.size();
    uint64_t chunks = 0;
    for (const EvPart &p : parts) chunks += (p.N + D - 1) / D;
    uint64_t base = 0;
    for (EvPart &p : parts) {
        const uint64_t ck = (p.N + D - 1) / D;
        if (n <= NS) {
            p.m = std::max<uint64_t>(1, std::min<uint64_t>(ck, 1 + (NS - n) * ck / std::max<uint64_t>(chunks, 1)));
            p.sbase = base;
            base += p.m;
        } else {
            p.m = NS;
            p.sbase = base % NS;
            base += ck;
        }
    }
}

// Networks whose synapses can run with implicit delay rings: Spiking or DS synapses (STDP needs
// per-step kernels), the reference's 32-bit rings, fused blocks.
static bool impl_capable(const bghip_net_desc &d)
{
    return (d.synapse_model == BGHIP_SYN_SPIKING || d.synapse_model == BGHIP_SYN_DS) && d.ring_bits == 32 &&
           d.block_steps != 1;
}

static void plan_net(bghip_net *n, Carve &c)
{
    const bghip_net_desc &d = n->desc;
    DevNet &dn = n->dn;
    size_t N = d.num_neurons, S = (size_t)N * d.max_synapses_per_neuron, W32 = d.ring_bits / 32;
    bool izh = d.neuron_model == BGHIP_NEURON_IZH;
    bool ds = d.synapse_model == BGHIP_SYN_DS || d.synapse_model == BGHIP_SYN_DYNSTDP;
    bool stdp = d.synapse_model == BGHIP_SYN_STDP || d.synapse_model == BGHIP_SYN_DYNSTDP;
    c.add(&dn.Vm, N); c.add(&dn.Vth, N); c.add(&dn.Vreset, N); c.add(&dn.I0, N); c.add(&dn.Inoise, N);
    c.add(&dn.C1, N); c.add(&dn.C2, N); c.add(&dn.Trefract, N); c.add(&dn.sp, N); c.add(&dn.refr, N);
    c.add(&dn.refr_steps, N); c.add(&dn.cnt, N); c.add(&dn.off, N);
    c.add(&dn.hist, N * (size_t)d.max_spikes); c.add(&dn.fired_step, N);
    c.add(&dn.mask[0], N); c.add(&dn.mask[1], N);
    if (izh) { c.add(&dn.A, N); c.add(&dn.B, N); c.add(&dn.C, N); c.add(&dn.D, N); c.add(&dn.u, N); c.add(&dn.C3, N); }
    c.add(&dn.meta, S); c.add(&dn.phase, S); c.add(&dn.psr, S); c.add(&dn.W, S); c.add(&dn.ring, S * W32);
    if (ds) { c.add(&dn.r, S); c.add(&dn.us, S); c.add(&dn.U, S); c.add(&dn.Dd, S); c.add(&dn.F, S); c.add(&dn.last, S); }
    if (stdp) {
        c.add(&dn.Apos, S); c.add(&dn.Aneg, S); c.add(&dn.gap, S); c.add(&dn.tauspost, S); c.add(&dn.tauspre, S);
        c.add(&dn.taupos, S); c.add(&dn.tauneg, S); c.add(&dn.Wex, S); c.add(&dn.mupos, S); c.add(&dn.muneg, S);
        c.add(&dn.useFD, S); c.add(&dn.stype, S); c.add(&dn.out_delay, S);
    }
    c.add(&dn.syn_count, N); c.add(&dn.slot_hi, N); c.add(&dn.in_map, S); c.add(&dn.in_begin, N);
    c.add(&dn.xloc, N); c.add(&dn.yloc, N); c.add(&dn.radii, N); c.add(&dn.rates, N); c.add(&dn.ntype, N);
    c.add(&dn.gbox, 5 * ((N + BG_GROWTH_TILE - 1) / BG_GROWTH_TILE));
    c.add(&dn.err, 1);
    c.add(&dn.in_mask, N); c.add(&dn.in_nisi, N); c.add(&dn.in_psr, N); c.add(&dn.in_r, N); c.add(&dn.in_u, N);
    c.add(&dn.in_last, N); c.add(&dn.in_ring, N); c.add(&dn.in_values, N); c.add(&dn.in_shift, N);
    if (stdp) {
        c.add(&dn.ev.count, 2 * BG_EV_SHARDS * BG_EV_PAD); c.add(&n->d_ev_off, 2 * (BG_EV_SHARDS + 1));
        c.add(&dn.ev.dest, N); c.add(&dn.ev.syn, S);
    }
    c.add(&n->d_self, 1);
    c.add(&dn.tbase, 1);
    c.add(&n->d_out_map, S); c.add(&n->d_out_begin, N); c.add(&n->d_out_count, N); c.add(&n->d_cursor, N);
    c.add(&n->d_stats, 4); c.add(&n->d_total, 1);
    c.add(&n->d_coh_used, 1);
    if (impl_capable(d)) {
        c.add(&dn.rec[0], N); c.add(&dn.rec[1], N); c.add(&dn.coh_step, BG_COH_MAX); c.add(&dn.Wd, S);
        c.add(&dn.lpt, N);
        if (ds) {
            const size_t C2N = (size_t)BG_COH_MAX * 2 * N;
            c.add(&dn.cst_r, C2N); c.add(&dn.cst_u, C2N); c.add(&dn.cst_last, C2N); c.add(&dn.cfac, 2 * C2N);
        }
    }
}

extern "C" int64_t bghip_net_footprint(const bghip_net_desc *desc)
{
    if (!desc_ok(desc)) return -1;
    bghip_net tmp;
    memset(&tmp, 0, sizeof tmp);
    tmp.desc = *desc;
    Carve c;
    plan_net(&tmp, c);
    return (int64_t)c.plan.off;
}

static int block_kmax(const bghip_net *n)
{
    int k = BG_KMAX;
    for (int t = 0; t < 4; t++) k = std::min(k, (int)n->desc.types.total_delay[t]);
    if (n->min_delay_uploaded > 0) k = std::min(k, n->min_delay_uploaded);
    if (n->desc.block_steps > 0) k = std::min(k, (int)n->desc.block_steps);
    if (n->desc.synapse_model == BGHIP_SYN_STDP || n->desc.synapse_model == BGHIP_SYN_DYNSTDP) k = 1;
    return std::max(k, 1);
}

static void set_kmax(bghip_net *n)
{
    n->kmax = block_kmax(n);
    n->dn.look = n->kmax;   // implicit rings: next-block window prepared by the source wavefronts
}

extern "C" int bghip_net_create(bghip_ctx *ctx, const bghip_net_desc *desc, bghip_net **out)
{
    if (!ctx || !out) return BGHIP_EINVAL;
    if (!desc_ok(desc)) return set_err(ctx, BGHIP_EINVAL, "invalid network description");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    bghip_net *n = new (std::nothrow) bghip_net();
    if (!n) return BGHIP_ENOMEM;
    memset((void *)n, 0, sizeof *n);
    n->ctx = ctx;
    n->desc = *desc;
    Carve c;
    plan_net(n, c);
    n->arena_bytes = c.plan.off;
    if (hipMalloc((void **)&n->arena, n->arena_bytes) != hipSuccess) {
        delete n;
        return set_err(ctx, BGHIP_ENOMEM, "hipMalloc of %zu bytes failed", c.plan.off);
    }
    for (auto &it : c.items) *it.first = n->arena + it.second;
    DevNet &dn = n->dn;
    dn.N = desc->num_neurons;
    dn.M = desc->max_synapses_per_neuron;
    dn.R = desc->ring_bits;
    dn.W32 = desc->ring_bits / 32;
    dn.max_spikes = desc->max_spikes;
    dn.dt = desc->dt;
    dn.seed = desc->noise_seed;
    dn.S = (uint64_t)dn.N * dn.M;
    for (int t = 0; t < 4; t++) {
        dn.decay[t] = desc->types.decay[t];
        dn.tU[t] = desc->types.U[t];
        dn.tD[t] = desc->types.D[t];
        dn.tF[t] = desc->types.F[t];
    }
    dn.ds_typed = 1;
    dn.stdp_typed = 1;
    for (int t = 0; t < 4; t++) dn.tdelay[t] = desc->types.total_delay[t];
    dn.self = n->d_self;
    dn.out_begin = n->d_out_begin;
    dn.out_count = n->d_out_count;
    dn.out_map = n->d_out_map;
    dn.ev.dest_off = n->d_ev_off;
    dn.ev.syn_off = n->d_ev_off ? n->d_ev_off + BG_EV_SHARDS + 1 : nullptr;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(n->arena);
        delete n;
        return set_err(ctx, BGHIP_EDEVICE, "hipStreamCreate failed");
    }
    n->h_total = nullptr;
    n->h_err = nullptr;
    n->h_coh_used = nullptr;
    if (hipHostMalloc((void **)&n->h_total, sizeof(uint64_t), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&n->h_err, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&n->h_coh_used, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        if (n->h_total) (void)hipHostFree(n->h_total);
        if (n->h_err) (void)hipHostFree(n->h_err);
        if (n->h_coh_used) (void)hipHostFree(n->h_coh_used);
        (void)hipStreamDestroy(n->stream);
        (void)hipFree(n->arena);
        delete n;
        return set_err(ctx, BGHIP_ENOMEM, "hipHostMalloc failed");
    }
    *n->h_total = 0;
    n->hist_range_known = true;
    HIP_TRY(ctx, hipEventCreate(&n->ev0));
    HIP_TRY(ctx, hipEventCreate(&n->ev1));
    HIP_TRY(ctx, hipEventCreateWithFlags(&n->ext_ev, hipEventDisableTiming));
    HIP_TRY(ctx, hipEventCreateWithFlags(&n->coh_ev, hipEventDisableTiming));
    // zero everything (empty synapse table, zero rings/counters), then the sentinels
    HIP_TRY(ctx, hipMemsetAsync(n->arena, 0, n->arena_bytes, n->stream));
    HIP_TRY(ctx, hipMemsetAsync(dn.fired_step, 0xFF, sizeof(uint64_t) * dn.N, n->stream));
    HIP_TRY(ctx, hipMemsetAsync(dn.hist, 0xFF, sizeof(uint64_t) * (size_t)dn.N * dn.max_spikes, n->stream));
    std::vector<uint32_t> doff, soff;
    if (n->d_ev_off) {
        if (!ev_offsets({{(uint64_t)dn.N, (uint64_t)dn.M, 0, BG_EV_SHARDS}}, BG_EV_SHARDS, doff, soff))
            return set_err(ctx, BGHIP_EUNSUPPORTED, "network too large for the event lists");
        doff.insert(doff.end(), soff.begin(), soff.end());
        HIP_TRY(ctx, hipMemcpyAsync(n->d_ev_off, doff.data(), sizeof(uint32_t) * doff.size(), hipMemcpyHostToDevice,
                                    n->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(n->stream));
    set_kmax(n);
    // (32-bit byte offsets address the slot arrays in the implicit kernels)
    dn.impl = impl_capable(*desc) && ctx->implicit && n->kmax >= 2 && dn.S < (1ull << 30) ? 1 : 0;
    *out = n;
    return BGHIP_OK;
}

static void drop_graph(bghip_net *n)
{
    if (n->g_exec) (void)hipGraphExecDestroy(n->g_exec);
    n->g_exec = nullptr;
}

static void batch_forget_net(bghip_batch *b, const bghip_net *n);

extern "C" void bghip_net_destroy(bghip_net *n)
{
    if (!n) return;
    (void)hipSetDevice(n->ctx->device);
    for (bghip_batch *b : n->ctx->batches) batch_forget_net(b, n);
    (void)hipStreamSynchronize(n->stream);
    drop_graph(n);
    (void)hipEventDestroy(n->ev0);
    (void)hipEventDestroy(n->ev1);
    (void)hipEventDestroy(n->ext_ev);
    if (n->coh_ev) (void)hipEventDestroy(n->coh_ev);
    (void)hipStreamDestroy(n->stream);
    (void)hipHostFree(n->h_coh_used);
    free(n->in_idx_host);
    free(n->in_mask_host);
    (void)hipHostFree(n->h_total);
    (void)hipHostFree(n->h_err);
    (void)hipFree(n->d_burst);
    (void)hipFree(n->d_shist);
    (void)hipFree(n->d_rates_hist);
    (void)hipFree(n->d_radii_hist);
    (void)hipFree(n->arena);
    delete n;
}

extern "C" uint64_t bghip_net_step(const bghip_net *n) { return n ? n->step : 0; }
extern "C" int bghip_net_implicit(const bghip_net *n) { return n ? n->dn.impl : 0; }
extern "C" int bghip_net_get_desc(const bghip_net *n, bghip_net_desc *out)
{
    if (!n || !out) return BGHIP_EINVAL;
    *out = n->desc;
    return BGHIP_OK;
}
extern "C" int bghip_recorder_info(const bghip_net *n, int64_t *n_burst, int64_t *n_spikes_hist, int32_t *num_epochs,
                                   int32_t *with_growth, int32_t *epoch)
{
    if (!n) return BGHIP_EINVAL;
    const bool on = n->rec_on;
    if (n_burst) *n_burst = on ? n->n_burst : 0;
    if (n_spikes_hist) *n_spikes_hist = on ? n->n_shist : 0;
    if (num_epochs) *num_epochs = on ? n->rec_epochs : 0;
    if (with_growth) *with_growth = on && n->rec_growth ? 1 : 0;
    if (epoch) *epoch = on ? n->rec_epoch : -1;
    return BGHIP_OK;
}

template <typename T>
static int h2d(bghip_net *n, T *dst, const T *src, size_t count)
{
    if (!src || !count) return BGHIP_OK;
    HIP_TRY(n->ctx, hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyHostToDevice, n->stream));
    return BGHIP_OK;
}
template <typename T>
static int d2h(bghip_net *n, T *dst, const T *src, size_t count)
{
    if (!dst || !count) return BGHIP_OK;
    HIP_TRY(n->ctx, hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost, n->stream));
    return BGHIP_OK;
}
#define CHK(x) do { int rc_ = (x); if (rc_ != BGHIP_OK) return rc_; } while (0)

// ------------------------------------------------------------------------------------------------
extern "C" int bghip_upload_neurons(bghip_net *n, const bghip_neurons *h)
{
    if (!n || !h) return BGHIP_EINVAL;
    net_begin(n);
    DevNet &dn = n->dn;
    size_t N = dn.N;
    if (!h->Vm || !h->Vthresh || !h->Vreset || !h->I0 || !h->Inoise || !h->C1 || !h->C2 || !h->Trefract)
        return set_err(n->ctx, BGHIP_EINVAL, "upload_neurons: missing LIF field");
    bool izh = n->desc.neuron_model == BGHIP_NEURON_IZH;
    if (izh && (!h->Aconst || !h->Bconst || !h->Cconst || !h->Dconst || !h->u || !h->C3))
        return set_err(n->ctx, BGHIP_EINVAL, "upload_neurons: missing Izhikevich field");
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    // the synapses' rings must keep their spikes while the neurons' histories are replaced
    if (n->dn.impl && net_total(n) > 0) CHK(to_explicit(n));
    // refractory length, AllLIFNeurons::fire (Neurons/AllLIFNeurons.cpp:73), computed on the host
    std::vector<int32_t> rs(N);
    for (size_t i = 0; i < N; i++) rs[i] = (int32_t)((double)(h->Trefract[i] / n->desc.dt) + 0.5);
    CHK(h2d(n, dn.Vm, h->Vm, N)); CHK(h2d(n, dn.Vth, h->Vthresh, N)); CHK(h2d(n, dn.Vreset, h->Vreset, N));
    CHK(h2d(n, dn.I0, h->I0, N)); CHK(h2d(n, dn.Inoise, h->Inoise, N)); CHK(h2d(n, dn.C1, h->C1, N));
    CHK(h2d(n, dn.C2, h->C2, N)); CHK(h2d(n, dn.Trefract, h->Trefract, N)); CHK(h2d(n, dn.refr_steps, rs.data(), N));
    CHK(h2d(n, dn.sp, h->summation_map, N)); CHK(h2d(n, dn.refr, h->nStepsInRefr, N));
    CHK(h2d(n, dn.cnt, h->spikeCount, N)); CHK(h2d(n, dn.off, h->spikeCountOffset, N));
    CHK(h2d(n, dn.hist, h->spike_history, N * (size_t)dn.max_spikes));
    // last firing step = the newest history entry (getSpikeHistory(i, -1)); the earliest step among
    // the counted entries bounds the recorder's bins at the next history update
    std::vector<uint64_t> last(N, BG_NO_SPIKE);
    uint64_t first_counted = n->step;
    if (h->spike_history && h->spikeCount && h->spikeCountOffset && dn.max_spikes > 0)
        for (size_t i = 0; i < N; i++) {
            const int ms = dn.max_spikes;
            const int k = (int)(((int64_t)h->spikeCount[i] + h->spikeCountOffset[i] + ms - 1) % ms);
            last[i] = h->spike_history[i * (size_t)ms + (size_t)k];
            for (int q = 0, idx = h->spikeCountOffset[i]; q < h->spikeCount[i] && q < ms; q++, idx++) {
                if (idx >= ms) idx = 0;
                first_counted = std::min<uint64_t>(first_counted, h->spike_history[i * (size_t)ms + (size_t)idx]);
            }
        }
    CHK(h2d(n, dn.fired_step, last.data(), N));
    n->hist_step0 = first_counted;
    n->hist_range_known = true;
    if (izh) {
        CHK(h2d(n, dn.A, h->Aconst, N)); CHK(h2d(n, dn.B, h->Bconst, N)); CHK(h2d(n, dn.C, h->Cconst, N));
        CHK(h2d(n, dn.D, h->Dconst, N)); CHK(h2d(n, dn.u, h->u, N)); CHK(h2d(n, dn.C3, h->C3, N));
    }
    HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
    return BGHIP_OK;
}

static int rebuild_incoming(bghip_net *n)
{
    DevNet &dn = n->dn;
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, n->stream, dn.syn_count, dn.in_begin, dn.N, n->d_total);
    if (dn.lpt) hipLaunchKernelGGL(k_lpt_order, dim3(1), dim3(1024), 0, n->stream, dn.syn_count, dn.lpt, dn.N);
    unsigned blocks = (unsigned)(((uint64_t)dn.N * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
    hipLaunchKernelGGL(k_fill_incoming, dim3(blocks), dim3(BG_BLOCK), 0, n->stream, dn);
    HIP_TRY(n->ctx, hipGetLastError());
    HIP_TRY(n->ctx, hipMemcpyAsync(n->h_total, n->d_total, sizeof(uint64_t), hipMemcpyDeviceToHost, n->stream));
    n->total_pending = true;
    n->out_valid = false;
    return BGHIP_OK;
}

// The synapse total, waiting for an index-map rebuild still in flight (growth updates do not wait
// for it, so the updates of many networks overlap on their streams).
static uint64_t net_total(bghip_net *n)
{
    if (n->total_pending) {
        (void)hipStreamSynchronize(n->stream);
        n->total = *n->h_total;
        n->total_pending = false;
    }
    return n->total;
}

// The outgoing map (Core/SynapseIndexMap.h: outgoingSynapseIndexMap/Begin/Count) from the incoming
// one: count per source, scan, place, then sort each source's segment ascending (the reference's
// map is stable in iSyn order). Built when first needed after the synapse table changed: by the
// per-step STDP path, which delivers spikes through it, or by an index-map download.
static int build_outgoing(bghip_net *n)
{
    DevNet &dn = n->dn;
    const size_t N = dn.N, T = net_total(n);
    if (n->out_valid) return BGHIP_OK;
    HIP_TRY(n->ctx, hipMemsetAsync(n->d_out_count, 0, 4 * N, n->stream));
    HIP_TRY(n->ctx, hipMemsetAsync(n->d_cursor, 0, 4 * N, n->stream));
    if (T > 0) {
        unsigned nb = (unsigned)((T + BG_BLOCK - 1) / BG_BLOCK);
        hipLaunchKernelGGL(k_count_outgoing, dim3(nb), dim3(BG_BLOCK), 0, n->stream, dn, dn.in_map, (uint64_t)T,
                           n->d_out_count);
        hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, n->stream, n->d_out_count, n->d_out_begin, dn.N,
                           n->d_total);
        hipLaunchKernelGGL(k_place_outgoing, dim3(nb), dim3(BG_BLOCK), 0, n->stream, dn, dn.in_map, (uint64_t)T,
                           n->d_out_begin, n->d_cursor, n->d_out_map);
        hipLaunchKernelGGL(k_sort_outgoing, dim3(dn.N), dim3(BG_BLOCK), 0, n->stream, n->d_out_begin,
                           n->d_out_count, n->d_out_map);
    } else {
        HIP_TRY(n->ctx, hipMemsetAsync(n->d_out_begin, 0, 4 * N, n->stream));
    }
    if (dn.stype && dn.S) {
        const unsigned nb = (unsigned)((dn.S + BG_BLOCK - 1) / BG_BLOCK);
        hipLaunchKernelGGL(k_slot_types, dim3(nb), dim3(BG_BLOCK), 0, n->stream, dn);
    }
    if (dn.out_delay && T > 0) {
        const unsigned nb = (unsigned)((T + BG_BLOCK - 1) / BG_BLOCK);
        hipLaunchKernelGGL(k_out_delays, dim3(nb), dim3(BG_BLOCK), 0, n->stream, dn, (uint64_t)T);
    }
    HIP_TRY(n->ctx, hipGetLastError());
    n->out_valid = true;
    return BGHIP_OK;
}

static inline uint32_t ring_get(const uint32_t *w, int b) { return (w[b >> 5] >> (b & 31)) & 1u; }

// Between the host's view of a device ring array (32-bit words, word w of slot s at w * S + s) and
// the byte planes STDP networks keep on the device (byte k of slot s at k * S + s; sring_word).
static void ring_words_to_planes(std::vector<uint32_t> &ring, size_t S, int W32)
{
    std::vector<uint32_t> out(ring.size(), 0u);
    uint8_t *b = (uint8_t *)out.data();
    for (int k = 0; k < 4 * W32; k++)
        for (size_t s = 0; s < S; s++) b[(size_t)k * S + s] = (uint8_t)(ring[(size_t)(k >> 2) * S + s] >> ((k & 3) * 8));
    ring.swap(out);
}
static void ring_planes_to_words(std::vector<uint32_t> &ring, size_t S, int W32)
{
    std::vector<uint32_t> out(ring.size(), 0u);
    const uint8_t *b = (const uint8_t *)ring.data();
    for (int k = 0; k < 4 * W32; k++)
        for (size_t s = 0; s < S; s++) out[(size_t)(k >> 2) * S + s] |= (uint32_t)b[(size_t)k * S + s] << ((k & 3) * 8);
    ring.swap(out);
}

// The per-type STDP table (DevNet::tp): a synapse of type t with parameters p either sets the
// entry of t or, if it differs from it in any bit, clears stdp_typed (the slot arrays stay the
// reference's and are what the kernels read from then on).
static void tp_merge(bghip_net *n, int t, const StdpParams &p)
{
    if (!n->tp_set[t]) {
        n->dn.tp[t] = p;
        n->tp_set[t] = true;
    } else if (memcmp(&n->dn.tp[t], &p, sizeof(StdpParams)) != 0) {
        n->dn.stdp_typed = 0;
    }
}
static StdpParams tp_of_ctx(const CreateCtx &c)
{
    StdpParams p;
    memset(&p, 0, sizeof p);
    p.Apos = c.Apos; p.Aneg = c.Aneg; p.gap = c.gap; p.tauspost = c.tauspost; p.tauspre = c.tauspre;
    p.taupos = c.taupos; p.tauneg = c.tauneg; p.Wex = c.Wex; p.mupos = c.mupos; p.muneg = c.muneg;
    p.useFD = (int32_t)(uint8_t)c.useFD;
    return p;
}

extern "C" int bghip_upload_synapses(bghip_net *n, const bghip_synapses *h, uint64_t step)
{
    if (!n || !h) return BGHIP_EINVAL;
    net_begin(n);
    DevNet &dn = n->dn;
    const int N = dn.N, M = dn.M, R = dn.R, W32 = dn.W32;
    const size_t S = dn.S;
    const int sk = n->desc.synapse_model;
    const bool ds = sk == BGHIP_SYN_DS || sk == BGHIP_SYN_DYNSTDP;
    const bool stdp = sk == BGHIP_SYN_STDP || sk == BGHIP_SYN_DYNSTDP;
    if (!h->in_use || !h->sourceNeuronIndex || !h->type || !h->W || !h->psr || !h->total_delay || !h->delayQueue ||
        !h->delayIdx || !h->synapse_counts)
        return set_err(n->ctx, BGHIP_EINVAL, "upload_synapses: missing field");
    if (ds && (!h->r || !h->u || !h->U || !h->D || !h->F || !h->lastSpike))
        return set_err(n->ctx, BGHIP_EINVAL, "upload_synapses: missing DS field");
    if (stdp && (!h->Apos || !h->Aneg || !h->STDPgap || !h->tauspost || !h->tauspre || !h->taupos || !h->tauneg ||
                 !h->Wex || !h->mupos || !h->muneg || !h->useFroemkeDanSTDP))
        return set_err(n->ctx, BGHIP_EINVAL, "upload_synapses: missing STDP field");
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    std::vector<uint32_t> meta(S, 0u), ring((size_t)W32 * S, 0u), cnt(N, 0u), hi(N, 0u);
    std::vector<uint8_t> phase(S, 0);
    int min_delay = 255;
    int typed = 1;
    if (stdp) {   // the uploaded slots replace every synapse: the table starts over
        for (int t = 0; t < 4; t++) n->tp_set[t] = false;
        dn.stdp_typed = 1;
    }
    for (int d = 0; d < N; d++) {
        for (int k = 0; k < M; k++) {
            size_t s = (size_t)d * M + k;
            if (!h->in_use[s]) continue;
            int src = h->sourceNeuronIndex[s], t = h->type[s], del = h->total_delay[s], di = h->delayIdx[s];
            if (src < 0 || src >= N || t < 0 || t > 3 || del < 1 || del >= R || del > 255 || di < 0 || di >= R)
                return set_err(n->ctx, BGHIP_EINVAL, "upload_synapses: slot %zu out of range", s);
            if (h->decay && memcmp(&h->decay[s], &dn.decay[t], 4) != 0)
                return set_err(n->ctx, BGHIP_EUNSUPPORTED, "upload_synapses: slot %zu decay differs from its type's", s);
            if (stdp && h->delayQueuePost && h->delayQueuePost[s] != 0)
                return set_err(n->ctx, BGHIP_EUNSUPPORTED, "upload_synapses: pending post-spike at a step boundary");
            if (stdp && h->delayIdxPost && (h->delayIdxPost[s] & 31) != (di & 31))
                return set_err(n->ctx, BGHIP_EUNSUPPORTED, "upload_synapses: pre/post ring phases differ");
            min_delay = std::min(min_delay, del);
            if (ds && (memcmp(&h->U[s], &dn.tU[t], 4) || memcmp(&h->D[s], &dn.tD[t], 4) || memcmp(&h->F[s], &dn.tF[t], 4)))
                typed = 0;
            if (stdp) {
                StdpParams p;
                memset(&p, 0, sizeof p);
                p.Apos = h->Apos[s]; p.Aneg = h->Aneg[s]; p.gap = h->STDPgap[s]; p.tauspost = h->tauspost[s];
                p.tauspre = h->tauspre[s]; p.taupos = h->taupos[s]; p.tauneg = h->tauneg[s]; p.Wex = h->Wex[s];
                p.mupos = h->mupos[s]; p.muneg = h->muneg[s]; p.useFD = (int32_t)(uint8_t)h->useFroemkeDanSTDP[s];
                tp_merge(n, t, p);
            }
            meta[s] = (uint32_t)src | ((uint32_t)t << BG_META_TYPE_SHIFT) | ((uint32_t)del << BG_META_DELAY_SHIFT) |
                      BG_META_IN_USE;
            // reference bit b (relative to delayIdx) -> absolute bit (b + phase) mod R,
            // phase = (step - delayIdx) mod R
            uint32_t ph = (uint32_t)((step - (uint64_t)di) & (uint64_t)(R - 1));
            phase[s] = (uint8_t)ph;
            const uint32_t *src_ring = &h->delayQueue[s * W32];
            for (int b = 0; b < R; b++)
                if (ring_get(src_ring, b)) {
                    int a = (b + (int)ph) & (R - 1);
                    ring[(size_t)(a >> 5) * S + s] |= 1u << (a & 31);
                }
            cnt[d]++;
            hi[d] = k + 1;
        }
        if (cnt[d] != h->synapse_counts[d])
            return set_err(n->ctx, BGHIP_EINVAL, "upload_synapses: synapse_counts[%d] disagrees with in_use", d);
    }
    n->min_delay_uploaded = cnt.empty() ? 0 : min_delay;
    dn.ds_typed = typed;
    // uploaded synapse state (rings, DS state) is taken as it is: the explicit kernels
    bool any = false;
    for (int d = 0; d < N && !any; d++) any = cnt[d] != 0;
    if (any) dn.impl = 0;
    if (stdp) ring_words_to_planes(ring, S, W32);
    CHK(h2d(n, dn.meta, meta.data(), S)); CHK(h2d(n, dn.phase, phase.data(), S));
    CHK(h2d(n, dn.ring, ring.data(), (size_t)W32 * S)); CHK(h2d(n, dn.psr, h->psr, S)); CHK(h2d(n, dn.W, h->W, S));
    CHK(h2d(n, dn.syn_count, cnt.data(), N)); CHK(h2d(n, dn.slot_hi, hi.data(), N));
    if (ds) {
        CHK(h2d(n, dn.r, h->r, S)); CHK(h2d(n, dn.us, h->u, S)); CHK(h2d(n, dn.U, h->U, S));
        CHK(h2d(n, dn.Dd, h->D, S)); CHK(h2d(n, dn.F, h->F, S)); CHK(h2d(n, dn.last, h->lastSpike, S));
    }
    if (stdp) {
        CHK(h2d(n, dn.Apos, h->Apos, S)); CHK(h2d(n, dn.Aneg, h->Aneg, S)); CHK(h2d(n, dn.gap, h->STDPgap, S));
        CHK(h2d(n, dn.tauspost, h->tauspost, S)); CHK(h2d(n, dn.tauspre, h->tauspre, S));
        CHK(h2d(n, dn.taupos, h->taupos, S)); CHK(h2d(n, dn.tauneg, h->tauneg, S)); CHK(h2d(n, dn.Wex, h->Wex, S));
        CHK(h2d(n, dn.mupos, h->mupos, S)); CHK(h2d(n, dn.muneg, h->muneg, S));
        CHK(h2d(n, dn.useFD, h->useFroemkeDanSTDP, S));
    }
    n->step = step;
    n->pending = false;
    set_kmax(n);
    CHK(rebuild_incoming(n));
    HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));   // the staging vectors above go out of scope
    return BGHIP_OK;
}

extern "C" int bghip_upload_layout(bghip_net *n, const float *xloc, const float *yloc, const int32_t *ntype)
{
    if (!n || !xloc || !yloc || !ntype) return BGHIP_EINVAL;
    net_begin(n);
    size_t N = n->dn.N;
    for (size_t i = 0; i < N; i++)
        if (ntype[i] != 1 && ntype[i] != 2) return set_err(n->ctx, BGHIP_EINVAL, "neuron type must be INH(1)/EXC(2)");
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    CHK(h2d(n, n->dn.xloc, xloc, N)); CHK(h2d(n, n->dn.yloc, yloc, N)); CHK(h2d(n, n->dn.ntype, ntype, N));
    HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
    n->ntype_host.assign(ntype, ntype + N);
    n->has_layout = true;
    return BGHIP_OK;
}

extern "C" int bghip_upload_growth(bghip_net *n, const bghip_growth *g, const float *radii, const float *rates)
{
    if (!n || !g || !radii) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    n->growth = *g;
    size_t N = n->dn.N;
    CHK(h2d(n, n->dn.radii, radii, N));
    if (rates) CHK(h2d(n, n->dn.rates, rates, N));
    HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
    n->has_growth = true;
    return BGHIP_OK;
}

// ------------------------------------------------------------------------------------------------
static void flush_pending(bghip_net *n)
{
    if (!n->pending) return;
    DevNet &dn = n->dn;
    unsigned blocks = (unsigned)((dn.S + BG_BLOCK - 1) / BG_BLOCK);
    if (blocks) hipLaunchKernelGGL(k_flush_block, dim3(blocks), dim3(BG_BLOCK), 0, n->stream, dn, n->last_t0, n->sel);
    n->pending = false;
}

// Implicit rings: writes the per-synapse rings and DS state the network's fire words and cohort
// states stand for at the current step into the slot arrays (for downloads and before the explicit
// kernels take over).
static int materialize(bghip_net *n)
{
    const DevNet &dn = n->dn;
    const unsigned blocks = (unsigned)((dn.S + BG_BLOCK - 1) / BG_BLOCK);
    if (blocks) hipLaunchKernelGGL(k_materialize, dim3(blocks), dim3(BG_BLOCK), 0, n->stream, dn, n->step, n->sel);
    HIP_TRY(n->ctx, hipGetLastError());
    return BGHIP_OK;
}

// Moves an implicit-ring network to the explicit kernels for the rest of its life (an upload of
// synapse or neuron state, per-synapse delays, more live cohorts than the meta bits hold, per-step
// blocks): its slot arrays then hold the complete rings, so the next block has nothing to pull.
static int to_explicit(bghip_net *n)
{
    if (!n->dn.impl) return BGHIP_OK;
    CHK(materialize(n));
    n->dn.impl = 0;
    n->pending = false;
    return BGHIP_OK;
}

// Issues a pass over the slots for the cohort ids still in use; its result is read at the next
// cohort allocation.
static int coh_scan(bghip_net *n)
{
    if (!n->dn.impl) return BGHIP_OK;
    HIP_TRY(n->ctx, hipMemsetAsync(n->d_coh_used, 0, sizeof(uint32_t), n->stream));
    const unsigned blocks = (unsigned)std::min<uint64_t>(1024, (n->dn.S + BG_BLOCK - 1) / BG_BLOCK + 1);
    hipLaunchKernelGGL(k_coh_usage, dim3(blocks), dim3(BG_BLOCK), 0, n->stream, n->dn, n->d_coh_used);
    HIP_TRY(n->ctx, hipGetLastError());
    HIP_TRY(n->ctx, hipMemcpyAsync(n->h_coh_used, n->d_coh_used, sizeof(uint32_t), hipMemcpyDeviceToHost, n->stream));
    HIP_TRY(n->ctx, hipEventRecord(n->coh_ev, n->stream));
    n->coh_scan_pending = true;
    n->coh_since = 0;
    return BGHIP_OK;
}

static int coh_collect(bghip_net *n)
{
    if (!n->coh_scan_pending) return BGHIP_OK;
    HIP_TRY(n->ctx, hipEventSynchronize(n->coh_ev));
    n->coh_alloc = *n->h_coh_used | n->coh_since;
    n->coh_scan_pending = false;
    n->dn.coh_mask = n->coh_alloc;
    return BGHIP_OK;
}

// Before a call that creates synapses on an implicit-ring network: a free cohort id for them, its
// creation step and createSynapse state (queued on the network's stream), in cc.coh. Ids of
// cohorts whose synapses are all gone come back through the usage scans; with every id live, the
// network moves to the explicit kernels.
static int coh_begin(bghip_net *n, CreateCtx &cc)
{
    cc.coh = 0;
    if (!n->dn.impl) return BGHIP_OK;
    CHK(coh_collect(n));
    const uint32_t all = (1u << BG_COH_MAX) - 1u;
    if (!(~n->coh_alloc & all)) {
        CHK(coh_scan(n));
        CHK(coh_collect(n));
        if (!(~n->coh_alloc & all)) return to_explicit(n);
    }
    const int c = __builtin_ffs((int)(~n->coh_alloc & all)) - 1;
    n->coh_alloc |= 1u << c;
    n->coh_since |= 1u << c;
    DevNet &dn = n->dn;
    dn.coh_mask = n->coh_alloc;
    dn.young_until = std::max<uint64_t>(dn.young_until, n->step + 64);
    const unsigned blocks = (unsigned)((2 * (uint64_t)dn.N + BG_BLOCK - 1) / BG_BLOCK);
    hipLaunchKernelGGL(k_coh_init, dim3(blocks), dim3(BG_BLOCK), 0, n->stream, dn, c, n->step, cc.r0, cc.u0);
    HIP_TRY(n->ctx, hipGetLastError());
    cc.coh = c;
    return BGHIP_OK;
}

// Captures BG_GRAPH_STEPS per-step STDP steps (offsets 0..S-1 from *tbase) and the tbase bump as
// one executable graph. Capture records only the kernels; work already queued on the stream stays
// ordered before each replay.
static int capture_stdp_graph(bghip_net *n)
{
    const DevNet &dn = n->dn;
    const int nk = n->desc.neuron_model, sk = n->desc.synapse_model;
    hipGraph_t g = nullptr;
    HIP_TRY(n->ctx, hipStreamBeginCapture(n->stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t i = 0; i < BG_GRAPH_STEPS; i++) launch_stdp_step(dn, n->d_self, n->stream, nk, sk, i);
    launch_step_bump(n->stream, dn.tbase, BG_GRAPH_STEPS);
    const hipError_t e = hipStreamEndCapture(n->stream, &g);
    if (e != hipSuccess) return set_err(n->ctx, BGHIP_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&n->g_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
        n->g_exec = nullptr;
        return set_err(n->ctx, BGHIP_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(e2));
    }
    memcpy((void *)&n->g_dn, (const void *)&dn, sizeof(DevNet));
    return BGHIP_OK;
}

extern "C" int bghip_advance(bghip_net *n, uint64_t n_steps)
{
    if (!n) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    DevNet &dn = n->dn;
    const int nk = n->desc.neuron_model, sk = n->desc.synapse_model;
    const bool per_step = n->kmax <= 1 || sk == BGHIP_SYN_STDP || sk == BGHIP_SYN_DYNSTDP;
    if (dn.impl && per_step) CHK(to_explicit(n));
    if (net_total(n) == 0) {
        // no synapses: neurons only, many steps per launch
        flush_pending(n);
        for (uint64_t done = 0; done < n_steps;) {
            const int k = (int)std::min<uint64_t>(BG_QUIET_STEPS, n_steps - done);
            launch_neuron_only(dn, n->stream, nk, n->step + done, k);
            done += (uint64_t)k;
            n->timed_launches += 1;
        }
        n->step += n_steps;
    } else if (per_step) {
        flush_pending(n);
        const bool stdp = sk == BGHIP_SYN_STDP || sk == BGHIP_SYN_DYNSTDP;
        if (stdp) CHK(build_outgoing(n));
        if (stdp && memcmp(&dn, &n->dn_shadow, sizeof(DevNet)) != 0) {
            // the event pass reads the descriptor from device memory; refresh it once nothing
            // queued on the stream can still be reading the old copy
            HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
            HIP_TRY(n->ctx, hipMemcpy(n->d_self, &dn, sizeof(DevNet), hipMemcpyHostToDevice));
            memcpy((void *)&n->dn_shadow, (const void *)&dn, sizeof(DevNet));
        }
        if (stdp) {
            // steps run at *tbase + offset: whole runs of BG_GRAPH_STEPS replay one captured graph
            // (its last node advances tbase), the rest is launched directly
            launch_step_set(n->stream, dn.tbase, n->step);
            uint64_t done = 0;
            if (n_steps >= BG_GRAPH_STEPS && n->ctx->graphs) {
                if (n->g_exec && memcmp(&dn, &n->g_dn, sizeof(DevNet)) != 0) drop_graph(n);
                if (!n->g_exec) CHK(capture_stdp_graph(n));
                for (; done + BG_GRAPH_STEPS <= n_steps; done += BG_GRAPH_STEPS)
                    HIP_TRY(n->ctx, hipGraphLaunch(n->g_exec, n->stream));
            }
            for (uint64_t i = 0; done + i < n_steps; i++)
                launch_stdp_step(dn, n->d_self, n->stream, nk, sk, (uint32_t)i);
            n->timed_launches += 4 * n_steps;
        }
        for (uint64_t i = 0; i < n_steps && !stdp; i++) {
            uint64_t t = n->step + i;
            {
                launch_neuron_step(dn, n->stream, nk, t);
                launch_synapse_step(dn, n->stream, sk, t);
                n->timed_launches += 2;
            }
        }
        n->step += n_steps;
    } else if (dn.impl) {
        // implicit rings: nothing to pull or flush between blocks or calls
        for (uint64_t done = 0; done < n_steps;) {
            const int K = (int)std::min<uint64_t>((uint64_t)n->kmax, n_steps - done);
            const uint64_t t0 = n->step + done;
            if (nk == BGHIP_NEURON_LIF) launch_fused_impl_lif(dn, n->stream, sk, n->kmax, t0, K, n->sel);
            else launch_fused_impl_izh(dn, n->stream, sk, n->kmax, t0, K, n->sel);
            n->sel ^= 1;
            n->last_t0 = t0;
            done += (uint64_t)K;
            n->timed_launches += 1;
        }
        n->step += n_steps;
        n->pending = false;
    } else {
        uint64_t done = 0;
        while (done < n_steps) {
            int K = (int)std::min<uint64_t>((uint64_t)n->kmax, n_steps - done);
            uint64_t t0 = n->step + done;
            int pull = n->pending ? 1 : 0;
            if (nk == BGHIP_NEURON_LIF)
                launch_fused_lif(dn, n->stream, sk, n->kmax, t0, K, pull, n->last_t0, n->sel);
            else
                launch_fused_izh(dn, n->stream, sk, n->kmax, t0, K, pull, n->last_t0, n->sel);
            n->sel ^= 1;
            n->last_t0 = t0;
            n->pending = true;
            done += (uint64_t)K;
            n->timed_launches += 1;
        }
        n->step += n_steps;
        flush_pending(n);
    }
    HIP_TRY(n->ctx, hipGetLastError());
    return BGHIP_OK;
}

// The error word copied behind everything queued on the stream into a pinned slot, then one wait.
static int read_error_bits(bghip_net *n, uint32_t *e)
{
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    HIP_TRY(n->ctx, hipMemcpyAsync(n->h_err, n->dn.err, 4, hipMemcpyDeviceToHost, n->stream));
    HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
    *e = *n->h_err;
    return BGHIP_OK;
}

extern "C" int bghip_wait(bghip_net *n)
{
    if (!n) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
    return BGHIP_OK;
}

extern "C" uint32_t bghip_net_error_bits(bghip_net *n)
{
    uint32_t e = 0;
    if (!n || read_error_bits(n, &e)) return 0;
    return e;
}

extern "C" int bghip_sync(bghip_net *n)
{
    if (!n) return BGHIP_EINVAL;
    uint32_t e = 0;
    CHK(read_error_bits(n, &e));
    if (e) return set_err(n->ctx, BGHIP_ESTATE, "device error bits 0x%x", e);
    return BGHIP_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// createSynapse's per-type constants (Synapses/AllSpikingSynapses.cpp, AllDSSynapses.cpp,
// AllSTDPSynapses.cpp createSynapse) for synapses created at the network's current step.
static CreateCtx create_ctx(const bghip_net *n)
{
    CreateCtx cc;
    memset(&cc, 0, sizeof cc);
    const bghip_synapse_types &ty = n->desc.types;
    cc.W_adj = (float)1.0e-8;
    for (int t = 0; t < 4; t++) {
        cc.total_delay[t] = ty.total_delay[t]; cc.U[t] = ty.U[t]; cc.D[t] = ty.D[t]; cc.F[t] = ty.F[t];
    }
    cc.u0 = ty.u0; cc.r0 = ty.r0;
    cc.Apos = ty.Apos; cc.Aneg = ty.Aneg; cc.gap = ty.STDPgap; cc.tauspost = ty.tauspost; cc.tauspre = ty.tauspre;
    cc.taupos = ty.taupos; cc.tauneg = ty.tauneg; cc.Wex = ty.Wex; cc.mupos = ty.mupos; cc.muneg = ty.muneg;
    cc.useFD = ty.useFroemkeDanSTDP;
    cc.syn_kind = n->desc.synapse_model;
    cc.step = n->step;
    return cc;
}

extern "C" int bghip_add_synapses(bghip_net *n, int64_t count, const int32_t *src, const int32_t *dest,
                                  const float *W)
{
    return bghip_add_synapses_ex(n, count, src, dest, W, nullptr, nullptr);
}

extern "C" int bghip_add_synapses_ex(bghip_net *n, int64_t count, const int32_t *src, const int32_t *dest,
                                     const float *W, const int32_t *total_delay, const float *Wex)
{
    if (!n || count < 0 || (count > 0 && (!src || !dest || !W))) return BGHIP_EINVAL;
    net_begin(n);
    if (!n->has_layout) return set_err(n->ctx, BGHIP_EINVAL, "add_synapses: layout not uploaded");
    const int N = n->dn.N;
    if (count == 0) return BGHIP_OK;
    // group by destination, keeping call order within each destination (counting sort)
    std::vector<int64_t> per(N + 1, 0);
    const int dmax = std::min(255, n->dn.R - 1);
    int dmin_new = 0;
    for (int64_t i = 0; i < count; i++) {
        if (src[i] < 0 || src[i] >= N || dest[i] < 0 || dest[i] >= N)
            return set_err(n->ctx, BGHIP_EINVAL, "add_synapses: neuron index out of range at entry %lld",
                           (long long)i);
        if (total_delay) {
            if (total_delay[i] < 1 || total_delay[i] > dmax)
                return set_err(n->ctx, BGHIP_EINVAL, "add_synapses: total delay %d at entry %lld outside 1..%d",
                               total_delay[i], (long long)i, dmax);
            dmin_new = dmin_new ? std::min(dmin_new, (int)total_delay[i]) : (int)total_delay[i];
        }
        per[dest[i] + 1]++;
    }
    std::vector<int32_t> dests;
    std::vector<int64_t> begin;
    for (int d = 0; d < N; d++) {
        if (per[d + 1]) { dests.push_back(d); begin.push_back(per[d]); }
        per[d + 1] += per[d];
    }
    begin.push_back(count);
    std::vector<int32_t> s2(count), d2(total_delay ? count : 0);
    std::vector<float> w2(count), x2(Wex ? count : 0);
    std::vector<int64_t> fill(per.begin(), per.end() - 1);
    for (int64_t i = 0; i < count; i++) {
        const int64_t k = fill[dest[i]]++;
        s2[k] = src[i];
        w2[k] = W[i];
        if (total_delay) d2[k] = total_delay[i];
        if (Wex) x2[k] = Wex[i];
    }
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    flush_pending(n);
    // per-synapse delays break the shared delivery times of a cohort: explicit rings
    if (total_delay) CHK(to_explicit(n));
    CreateCtx cc = create_ctx(n);
    CHK(coh_begin(n, cc));
    if (cc.syn_kind == SK_STDP || cc.syn_kind == SK_DYNSTDP) {
        StdpParams p = tp_of_ctx(cc);
        if (!Wex) {
            for (int t = 0; t < 4; t++) tp_merge(n, t, p);
        } else {
            // createSynapse's values with the caller's Wex, under the type the device gives the synapse
            // (syn_type_of: II 0, IE 1, EI 2, EE 3 from the two neurons' types, Layouts/Layout.cpp:129-141)
            for (int64_t i = 0; i < count; i++) {
                const int ts = n->ntype_host[src[i]], td = n->ntype_host[dest[i]];
                p.Wex = Wex[i];
                tp_merge(n, ts == 1 ? (td == 1 ? 0 : 1) : (td == 1 ? 2 : 3), p);
            }
        }
    }
    const int nd = (int)dests.size();
    char *buf = nullptr;
    const size_t o_dest = 0, o_begin = align256(sizeof(int32_t) * nd),
                 o_src = o_begin + align256(sizeof(int64_t) * (nd + 1)), o_w = o_src + align256(sizeof(int32_t) * count),
                 o_d = o_w + align256(sizeof(float) * count), o_x = o_d + align256(sizeof(int32_t) * d2.size()),
                 bytes = o_x + sizeof(float) * x2.size() + 4;
    if (hipMalloc((void **)&buf, bytes) != hipSuccess) return set_err(n->ctx, BGHIP_ENOMEM, "add_synapses: staging");
    int rc = BGHIP_OK;
    if (hipMemcpyAsync(buf + o_dest, dests.data(), sizeof(int32_t) * nd, hipMemcpyHostToDevice, n->stream) ||
        hipMemcpyAsync(buf + o_begin, begin.data(), sizeof(int64_t) * (nd + 1), hipMemcpyHostToDevice, n->stream) ||
        hipMemcpyAsync(buf + o_src, s2.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice, n->stream) ||
        hipMemcpyAsync(buf + o_w, w2.data(), sizeof(float) * count, hipMemcpyHostToDevice, n->stream) ||
        (total_delay &&
         hipMemcpyAsync(buf + o_d, d2.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice, n->stream)) ||
        (Wex && hipMemcpyAsync(buf + o_x, x2.data(), sizeof(float) * count, hipMemcpyHostToDevice, n->stream))) {
        rc = set_err(n->ctx, BGHIP_EDEVICE, "add_synapses: upload");
    } else {
        const unsigned grid = (unsigned)(((int64_t)nd * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
        hipLaunchKernelGGL(k_add_synapses, dim3(grid), dim3(BG_BLOCK), 0, n->stream, n->dn, cc,
                           (const int32_t *)(buf + o_dest), (const int64_t *)(buf + o_begin),
                           (const int32_t *)(buf + o_src), (const float *)(buf + o_w),
                           total_delay ? (const int32_t *)(buf + o_d) : nullptr,
                           Wex ? (const float *)(buf + o_x) : nullptr, nd);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(n->stream) != hipSuccess)
            rc = set_err(n->ctx, BGHIP_EDEVICE, "add_synapses: kernel");
    }
    (void)hipFree(buf);
    if (rc) return rc;
    if (dmin_new) {   // per-synapse delays bound the fused block length like uploaded ones
        n->min_delay_uploaded = n->min_delay_uploaded ? std::min(n->min_delay_uploaded, dmin_new) : dmin_new;
        set_kmax(n);
    }
    CHK(rebuild_incoming(n));
    CHK(coh_scan(n));
    return BGHIP_OK;
}

static int apply_weights(bghip_net *n, const float *Wd, bghip_growth_stats *stats);

extern "C" int bghip_update_connections(bghip_net *n, bghip_growth_stats *stats)
{
    if (!n) return BGHIP_EINVAL;
    net_begin(n);
    if (!n->has_layout || !n->has_growth) return set_err(n->ctx, BGHIP_EINVAL, "layout/growth state not uploaded");
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    DevNet &dn = n->dn;
    flush_pending(n);
    const bghip_growth &g = n->growth;
    unsigned nb = (unsigned)((dn.N + BG_BLOCK - 1) / BG_BLOCK);
    float c = g.epochDuration * g.rho;
    hipLaunchKernelGGL(k_growth_radii, dim3(nb), dim3(BG_BLOCK), 0, n->stream, dn, g.epochDuration, g.epsilon,
                       g.maxRate, g.beta, c);
    return apply_weights(n, nullptr, stats);
}

// ConnGrowth::updateSynapsesWeights (Connections/ConnGrowth.cpp:388-463) with W(s, d) = the overlap
// areas (Wd null) or a device copy of the caller's matrix, then the incoming map.
static int apply_weights(bghip_net *n, const float *Wd, bghip_growth_stats *stats)
{
    DevNet &dn = n->dn;
    CreateCtx cc = create_ctx(n);
    CHK(coh_begin(n, cc));
    if (cc.syn_kind == SK_STDP || cc.syn_kind == SK_DYNSTDP)
        for (int t = 0; t < 4; t++) tp_merge(n, t, tp_of_ctx(cc));
    HIP_TRY(n->ctx, hipMemsetAsync(n->d_stats, 0, 16, n->stream));
    if (dn.M > 0) {
        unsigned grid = (unsigned)std::min(dn.N, 4096);
        if (!Wd)
            hipLaunchKernelGGL(k_growth_tiles, dim3((unsigned)((dn.N + BG_GROWTH_TILE - 1) / BG_GROWTH_TILE)),
                               dim3(BG_BLOCK), 0, n->stream, dn);
        hipLaunchKernelGGL(k_growth_dest, dim3(grid), dim3(BG_BLOCK), 0, n->stream, dn, cc, n->d_stats, Wd);
    }
    HIP_TRY(n->ctx, hipGetLastError());
    CHK(rebuild_incoming(n));
    CHK(coh_scan(n));
    if (stats) {
        uint32_t st[4];
        HIP_TRY(n->ctx, hipMemcpyAsync(st, n->d_stats, 16, hipMemcpyDeviceToHost, n->stream));
        stats->total_synapses = (int64_t)net_total(n);   // waits for the stream
        HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
        stats->created = st[0];
        stats->erased = st[1];
    }
    return BGHIP_OK;
}

// The caller's matrix W(s, d) (row-major, s the source) in device memory at Wd is transposed in
// place, so each destination's workgroup reads its row contiguously, then applied.
static int apply_caller_weights(bghip_net *n, float *Wd, const float *W, bghip_growth_stats *stats)
{
    const int N = n->dn.N;
    const size_t bytes = sizeof(float) * (size_t)N * (size_t)N;
    if (hipMemcpyAsync(Wd, W, bytes, hipMemcpyHostToDevice, n->stream) != hipSuccess)
        return set_err(n->ctx, BGHIP_EDEVICE, "weight matrix upload failed");
    const unsigned T = (unsigned)((N + BG_TT - 1) / BG_TT);
    hipLaunchKernelGGL(k_transpose_square, dim3(T, T), dim3(BG_BLOCK), 0, n->stream, Wd, N);
    HIP_TRY(n->ctx, hipGetLastError());
    return apply_weights(n, Wd, stats);
}

extern "C" int bghip_update_weights(bghip_net *n, const float *W, bghip_growth_stats *stats)
{
    if (!n || !W) return BGHIP_EINVAL;
    net_begin(n);
    if (!n->has_layout) return set_err(n->ctx, BGHIP_EINVAL, "layout not uploaded");
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    flush_pending(n);
    const size_t bytes = sizeof(float) * (size_t)n->dn.N * (size_t)n->dn.N;
    float *Wd = nullptr;
    HIP_TRY(n->ctx, hipMalloc((void **)&Wd, std::max<size_t>(bytes, 4)));
    int rc = apply_caller_weights(n, Wd, W, stats);
    // the matrix is freed once the kernels reading it are done
    const hipError_t e = hipStreamSynchronize(n->stream);
    (void)hipFree(Wd);
    if (rc == BGHIP_OK && e != hipSuccess) rc = set_err(n->ctx, BGHIP_EDEVICE, "%s", hipGetErrorString(e));
    return rc;
}

extern "C" int bghip_update_weights_static(bghip_net *n, const float *W, bghip_growth_stats *stats)
{
    if (!n || !W) return BGHIP_EINVAL;
    net_begin(n);
    if (!n->has_layout) return set_err(n->ctx, BGHIP_EINVAL, "layout not uploaded");
    const size_t bytes = sizeof(float) * (size_t)n->dn.N * (size_t)n->dn.N;
    if (bytes > BG_SCRATCH_BYTES)
        return set_err(n->ctx, BGHIP_EUNSUPPORTED,
                       "update_weights_static: a %d x %d weight matrix (%zu bytes) exceeds the static scratch of "
                       "%u bytes (at most %d neurons)", n->dn.N, n->dn.N, bytes, (unsigned)BG_SCRATCH_BYTES,
                       BG_SCRATCH_MAX_N);
    float *Wd = (float *)n->ctx->scratch;
    if (!Wd) return set_err(n->ctx, BGHIP_EDEVICE, "update_weights_static: static scratch address unavailable");
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    flush_pending(n);
    int rc = apply_caller_weights(n, Wd, W, stats);
    // the scratch is free for the next call (on any network of this device) once this returns
    const hipError_t e = hipStreamSynchronize(n->stream);
    if (rc == BGHIP_OK && e != hipSuccess) rc = set_err(n->ctx, BGHIP_EDEVICE, "%s", hipGetErrorString(e));
    return rc;
}

extern "C" int bghip_update_history(bghip_net *n)
{
    if (!n) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    const unsigned nb = (unsigned)((n->dn.N + BG_BLOCK - 1) / BG_BLOCK);
    float *rates_row = nullptr, *radii_row = nullptr;
    if (n->rec_on) {
        // bins the epoch's spikes can fall in: steps [hist_step0, step)
        int64_t b_lo = 0, s_lo = 0;
        int nb_loc = 0, ns_loc = 0;
        if (n->hist_range_known && n->step > n->hist_step0) {
            const double dt = (double)n->dn.dt;
            const double x0 = (double)n->hist_step0 * dt, x1 = (double)(n->step - 1) * dt;
            b_lo = (int64_t)(int)x0;
            s_lo = (int64_t)(int)(x0 * 100);
            const int64_t bw = (int64_t)(int)x1 - b_lo + 1, sw = (int64_t)(int)(x1 * 100) - s_lo + 1;
            if (bw > 0 && sw > 0 && bw + sw <= 8192) { nb_loc = (int)bw; ns_loc = (int)sw; }
        }
        hipLaunchKernelGGL(k_record_spikes, dim3(nb), dim3(BG_BLOCK), sizeof(uint32_t) * (nb_loc + ns_loc), n->stream,
                           n->dn, n->d_burst, n->n_burst, n->d_shist, n->n_shist, b_lo, nb_loc, s_lo, ns_loc);
        const int e = n->rec_epoch + 1;
        if (n->rec_growth && e <= n->rec_epochs) {
            rates_row = n->d_rates_hist + (size_t)e * n->dn.N;
            radii_row = n->d_radii_hist + (size_t)e * n->dn.N;
        }
        n->rec_epoch = e;
    }
    hipLaunchKernelGGL(k_update_history, dim3(nb), dim3(BG_BLOCK), 0, n->stream, n->dn, n->growth.minRadius,
                       n->has_growth ? 1 : 0, rates_row, radii_row);
    HIP_TRY(n->ctx, hipGetLastError());
    n->hist_step0 = n->step;
    n->hist_range_known = true;
    return BGHIP_OK;
}

extern "C" int bghip_recorder_enable(bghip_net *n, int64_t n_burst, int64_t n_spikes_hist, int32_t num_epochs,
                                     int32_t with_growth, const float *rates0, const float *radii0)
{
    if (!n || n_burst < 0 || n_spikes_hist < 0 || num_epochs < 0) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
    (void)hipFree(n->d_burst); (void)hipFree(n->d_shist);
    (void)hipFree(n->d_rates_hist); (void)hipFree(n->d_radii_hist);
    n->d_burst = n->d_shist = nullptr;
    n->d_rates_hist = n->d_radii_hist = nullptr;
    const size_t N = n->dn.N, rows = (size_t)num_epochs + 1;
    if (hipMalloc((void **)&n->d_burst, sizeof(uint32_t) * (size_t)std::max<int64_t>(n_burst, 1)) != hipSuccess ||
        hipMalloc((void **)&n->d_shist, sizeof(uint32_t) * (size_t)std::max<int64_t>(n_spikes_hist, 1)) != hipSuccess)
        return set_err(n->ctx, BGHIP_ENOMEM, "recorder: hipMalloc failed");
    HIP_TRY(n->ctx, hipMemsetAsync(n->d_burst, 0, sizeof(uint32_t) * (size_t)n_burst, n->stream));
    HIP_TRY(n->ctx, hipMemsetAsync(n->d_shist, 0, sizeof(uint32_t) * (size_t)n_spikes_hist, n->stream));
    if (with_growth) {
        if (hipMalloc((void **)&n->d_rates_hist, sizeof(float) * rows * N) != hipSuccess ||
            hipMalloc((void **)&n->d_radii_hist, sizeof(float) * rows * N) != hipSuccess)
            return set_err(n->ctx, BGHIP_ENOMEM, "recorder: hipMalloc failed");
        HIP_TRY(n->ctx, hipMemsetAsync(n->d_rates_hist, 0, sizeof(float) * rows * N, n->stream));
        HIP_TRY(n->ctx, hipMemsetAsync(n->d_radii_hist, 0, sizeof(float) * rows * N, n->stream));
        CHK(h2d(n, n->d_rates_hist, rates0, N));
        CHK(h2d(n, n->d_radii_hist, radii0, N));
    }
    HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
    n->n_burst = n_burst;
    n->n_shist = n_spikes_hist;
    n->rec_epochs = num_epochs;
    n->rec_epoch = 0;
    n->rec_growth = with_growth != 0;
    n->rec_on = true;
    return BGHIP_OK;
}

static void counts_to_float(const std::vector<uint32_t> &c, float *out)
{
    for (size_t k = 0; k < c.size(); k++) out[k] = (float)std::min<uint32_t>(c[k], 1u << 24);
}

extern "C" int bghip_recorder_download(bghip_net *n, float *burstinessHist, float *spikesHistory,
                                       float *ratesHistory, float *radiiHistory)
{
    if (!n || !n->rec_on) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    std::vector<uint32_t> b((size_t)n->n_burst), sh((size_t)n->n_shist);
    CHK(d2h(n, b.data(), n->d_burst, b.size()));
    CHK(d2h(n, sh.data(), n->d_shist, sh.size()));
    const size_t cells = ((size_t)n->rec_epochs + 1) * n->dn.N;
    if (n->rec_growth) {
        CHK(d2h(n, ratesHistory, n->d_rates_hist, cells));
        CHK(d2h(n, radiiHistory, n->d_radii_hist, cells));
    }
    CHK(bghip_sync(n));
    if (burstinessHist) counts_to_float(b, burstinessHist);
    if (spikesHistory) counts_to_float(sh, spikesHistory);
    return BGHIP_OK;
}

// ------------------------------------------------------------------------------------------------
extern "C" int bghip_download_neurons(bghip_net *n, bghip_neurons *h)
{
    if (!n || !h) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    DevNet &dn = n->dn;
    size_t N = dn.N;
    CHK(d2h(n, h->Vm, dn.Vm, N)); CHK(d2h(n, h->Vthresh, dn.Vth, N)); CHK(d2h(n, h->Vreset, dn.Vreset, N));
    CHK(d2h(n, h->I0, dn.I0, N)); CHK(d2h(n, h->Inoise, dn.Inoise, N)); CHK(d2h(n, h->C1, dn.C1, N));
    CHK(d2h(n, h->C2, dn.C2, N)); CHK(d2h(n, h->Trefract, dn.Trefract, N)); CHK(d2h(n, h->summation_map, dn.sp, N));
    CHK(d2h(n, h->nStepsInRefr, dn.refr, N)); CHK(d2h(n, h->spikeCount, dn.cnt, N));
    CHK(d2h(n, h->spikeCountOffset, dn.off, N)); CHK(d2h(n, h->spike_history, dn.hist, N * (size_t)dn.max_spikes));
    if (n->desc.neuron_model == BGHIP_NEURON_IZH) {
        CHK(d2h(n, h->Aconst, dn.A, N)); CHK(d2h(n, h->Bconst, dn.B, N)); CHK(d2h(n, h->Cconst, dn.C, N));
        CHK(d2h(n, h->Dconst, dn.D, N)); CHK(d2h(n, h->u, dn.u, N)); CHK(d2h(n, h->C3, dn.C3, N));
    }
    return bghip_sync(n);
}

extern "C" int bghip_download_spikes(bghip_net *n, int32_t *counts, int32_t *offsets, uint64_t *history)
{
    if (!n) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    const DevNet &dn = n->dn;
    const size_t N = dn.N;
    if (counts) CHK(d2h(n, counts, dn.cnt, N));
    if (offsets) CHK(d2h(n, offsets, dn.off, N));
    if (history) CHK(d2h(n, history, dn.hist, N * (size_t)dn.max_spikes));
    return bghip_sync(n);
}

extern "C" int bghip_clear_spike_counts(bghip_net *n)
{
    if (!n) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    // k_update_history without the radius clamp and history rows is exactly clearSpikeCounts
    const unsigned nb = (unsigned)((n->dn.N + BG_BLOCK - 1) / BG_BLOCK);
    hipLaunchKernelGGL(k_update_history, dim3(nb), dim3(BG_BLOCK), 0, n->stream, n->dn, 0.0f, 0, nullptr, nullptr);
    HIP_TRY(n->ctx, hipGetLastError());
    n->hist_step0 = n->step;   // the recorder's next compile starts after the cleared spikes
    n->hist_range_known = true;
    return BGHIP_OK;
}

extern "C" int64_t bghip_total_synapses(bghip_net *n) { return n ? (int64_t)net_total(n) : -1; }

extern "C" int bghip_download_synapses(bghip_net *n, bghip_synapses *h)
{
    if (!n || !h) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    if (n->dn.impl) CHK(materialize(n));
    DevNet &dn = n->dn;
    const int N = dn.N, M = dn.M, R = dn.R, W32 = dn.W32;
    const size_t S = dn.S;
    const int sk = n->desc.synapse_model;
    const bool ds = sk == BGHIP_SYN_DS || sk == BGHIP_SYN_DYNSTDP;
    const bool stdp = sk == BGHIP_SYN_STDP || sk == BGHIP_SYN_DYNSTDP;
    std::vector<uint32_t> meta(S), ring((size_t)W32 * S);
    std::vector<uint8_t> phase(S);
    CHK(d2h(n, meta.data(), dn.meta, S)); CHK(d2h(n, phase.data(), dn.phase, S));
    CHK(d2h(n, ring.data(), dn.ring, (size_t)W32 * S));
    CHK(d2h(n, h->psr, dn.psr, S)); CHK(d2h(n, h->W, dn.W, S)); CHK(d2h(n, h->synapse_counts, dn.syn_count, N));
    if (ds) {
        CHK(d2h(n, h->r, dn.r, S)); CHK(d2h(n, h->u, dn.us, S)); CHK(d2h(n, h->U, dn.U, S));
        CHK(d2h(n, h->D, dn.Dd, S)); CHK(d2h(n, h->F, dn.F, S)); CHK(d2h(n, h->lastSpike, dn.last, S));
    }
    if (stdp) {
        CHK(d2h(n, h->Apos, dn.Apos, S)); CHK(d2h(n, h->Aneg, dn.Aneg, S)); CHK(d2h(n, h->STDPgap, dn.gap, S));
        CHK(d2h(n, h->tauspost, dn.tauspost, S)); CHK(d2h(n, h->tauspre, dn.tauspre, S));
        CHK(d2h(n, h->taupos, dn.taupos, S)); CHK(d2h(n, h->tauneg, dn.tauneg, S)); CHK(d2h(n, h->Wex, dn.Wex, S));
        CHK(d2h(n, h->mupos, dn.mupos, S)); CHK(d2h(n, h->muneg, dn.muneg, S));
        CHK(d2h(n, h->useFroemkeDanSTDP, dn.useFD, S));
    }
    CHK(bghip_sync(n));
    if (stdp) ring_planes_to_words(ring, S, W32);
    const uint64_t step = n->step;
    for (size_t s = 0; s < S; s++) {
        uint32_t m = meta[s];
        bool use = (m & BG_META_IN_USE) != 0;
        int type = (m >> BG_META_TYPE_SHIFT) & 3;
        uint32_t ph = phase[s];
        if (h->in_use) h->in_use[s] = use ? 1 : 0;
        if (h->sourceNeuronIndex) h->sourceNeuronIndex[s] = use ? (int32_t)(m & BG_META_SRC_MASK) : 0;
        if (h->destNeuronIndex) h->destNeuronIndex[s] = use ? (int32_t)(s / (size_t)M) : 0;
        if (h->type) h->type[s] = use ? type : -1;
        if (h->total_delay) h->total_delay[s] = use ? (int32_t)((m >> BG_META_DELAY_SHIFT) & 0xFF) : 0;
        if (h->decay) h->decay[s] = use ? dn.decay[type] : 0.0f;
        int di = (int)((step - (uint64_t)ph) & (uint64_t)(R - 1));
        if (h->delayIdx) h->delayIdx[s] = use ? di : 0;
        if (h->delayQueue) {
            uint32_t *dst = &h->delayQueue[s * W32];
            for (int w = 0; w < W32; w++) dst[w] = 0u;
            if (use)
                for (int a = 0; a < R; a++)
                    if ((ring[(size_t)(a >> 5) * S + s] >> (a & 31)) & 1u) {
                        int b = (a - (int)ph) & (R - 1);
                        dst[b >> 5] |= 1u << (b & 31);
                    }
        }
        if (stdp) {
            if (h->delayQueuePost) h->delayQueuePost[s] = 0u;
            if (h->delayIdxPost) h->delayIdxPost[s] = use ? (int32_t)((step - (uint64_t)ph) & 31ull) : 0;
        }
    }
    return BGHIP_OK;
}

extern "C" int bghip_download_index_map(bghip_net *n, bghip_index_map *h)
{
    if (!n || !h) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    DevNet &dn = n->dn;
    size_t N = dn.N, T = net_total(n);
    CHK(build_outgoing(n));
    std::vector<uint32_t> cnt(N);
    CHK(d2h(n, h->incomingSynapseIndexMap, dn.in_map, T)); CHK(d2h(n, h->incomingSynapseBegin, dn.in_begin, N));
    CHK(d2h(n, cnt.data(), dn.syn_count, N));
    CHK(d2h(n, h->outgoingSynapseIndexMap, n->d_out_map, T)); CHK(d2h(n, h->outgoingSynapseBegin, n->d_out_begin, N));
    CHK(d2h(n, h->outgoingSynapseCount, n->d_out_count, N));
    CHK(bghip_sync(n));
    if (h->incomingSynapseCount) memcpy(h->incomingSynapseCount, cnt.data(), 4 * N);
    return BGHIP_OK;
}

extern "C" int bghip_download_growth(bghip_net *n, float *radii, float *rates)
{
    if (!n) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    CHK(d2h(n, radii, n->dn.radii, n->dn.N));
    CHK(d2h(n, rates, n->dn.rates, n->dn.N));
    return bghip_sync(n);
}

// ------------------------------------------------------------------------------------------------
// Stimulus input. The input synapse constants are those of a freshly created AllDSSynapses
// synapse of type EE (index 0) or EI (index 1): Synapses/AllSpikingSynapses.cpp:191-233 and
// Synapses/AllDSSynapses.cpp:180-218, decay on the host C library (AllSpikingSynapses.cpp:344-354).
extern "C" int bghip_upload_input(bghip_net *n, const bghip_input *in, const bghip_input_state *st)
{
    if (!n) return BGHIP_EINVAL;
    net_begin(n);
    DevNet &dn = n->dn;
    const int N = dn.N;
    const float dt = dn.dt;
    if (!in || in->kind == BGHIP_INPUT_NONE) {
        dn.in_kind = BG_INPUT_NONE;
        return BGHIP_OK;
    }
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    if (in->kind == BGHIP_INPUT_POISSON) {
        if (!in->neuron_type) return set_err(n->ctx, BGHIP_EINVAL, "upload_input: Poisson input needs neuron_type");
        if (!(in->fr_mean > 0.0f)) return set_err(n->ctx, BGHIP_EINVAL, "upload_input: fr_mean must be positive");
        static const float tau[2] = {(float)3e-3, (float)3e-3}, delay[2] = {(float)1.5e-3, (float)0.8e-3};
        static const float U[2] = {(float)0.5, (float)0.05}, D[2] = {(float)1.1, (float)0.125};
        static const float F[2] = {(float)0.05, (float)1.2};
        for (int k = 0; k < 2; k++) {
            dn.in_decay[k] = expf(-dt / tau[k]);
            dn.in_delay[k] = (int)(delay[k] / dt) + 1;
            dn.in_U[k] = U[k]; dn.in_D[k] = D[k]; dn.in_F[k] = F[k];
            if (dn.in_delay[k] < 1 || dn.in_delay[k] >= 32)
                return set_err(n->ctx, BGHIP_EUNSUPPORTED, "upload_input: input delay does not fit its ring");
        }
        const float fr = in->fr_mean / 1000;   // per ms (Inputs/SInputPoisson.cpp:55-57)
        dn.in_lambda = 1 / fr;
        dn.in_W = in->weight * (float)1.0e-8;  // AllSynapses::SYNAPSE_STRENGTH_ADJUSTMENT
        std::vector<uint8_t> mask(N);
        for (int i = 0; i < N; i++) {
            if (in->neuron_type[i] != 1 && in->neuron_type[i] != 2)
                return set_err(n->ctx, BGHIP_EINVAL, "upload_input: neuron type must be INH(1)/EXC(2)");
            const bool on = in->mask ? in->mask[i] != 0 : true;
            mask[i] = on ? (uint8_t)(1 + (in->neuron_type[i] == 1 ? 1 : 0)) : 0;
        }
        std::vector<int32_t> nisi(N, 0);
        std::vector<float> psr(N, 0.0f), r(N, 1.0f), u(N, (float)0.4);
        std::vector<uint64_t> last(N, BG_NO_SPIKE);
        std::vector<uint32_t> ring(N, 0u);
        uint32_t phase = (uint32_t)(n->step & 31u);
        if (st) {
            if (st->nISIs) std::copy(st->nISIs, st->nISIs + N, nisi.begin());
            if (st->psr) std::copy(st->psr, st->psr + N, psr.begin());
            if (st->r) std::copy(st->r, st->r + N, r.begin());
            if (st->u) std::copy(st->u, st->u + N, u.begin());
            if (st->lastSpike) std::copy(st->lastSpike, st->lastSpike + N, last.begin());
            if (st->delayQueue) {
                // every neuron with input advances its input synapse each step, so they share one
                // delayIdx; the synapses of neurons without input keep theirs
                if (!st->delayIdx) return set_err(n->ctx, BGHIP_EINVAL, "upload_input: delayQueue needs delayIdx");
                int di = -1;
                for (int i = 0; i < N; i++) {
                    if (st->delayIdx[i] < 0 || st->delayIdx[i] >= 32)
                        return set_err(n->ctx, BGHIP_EINVAL, "upload_input: delayIdx out of range");
                    if (!mask[i]) continue;
                    if (di < 0) di = st->delayIdx[i];
                    if (st->delayIdx[i] != di)
                        return set_err(n->ctx, BGHIP_EUNSUPPORTED, "upload_input: input rings of different phases");
                }
                if (di < 0) di = 0;
                phase = (uint32_t)((n->step - (uint64_t)di) & 31u);
                // reference bit b (read at delayIdx when the step is `step`) -> absolute bit b + phase
                for (int i = 0; i < N; i++) {
                    const uint32_t q = st->delayQueue[i];
                    ring[i] = mask[i] && phase ? (q << phase) | (q >> (32u - phase)) : q;
                }
            }
        }
        n->in_phase = phase;
        if (!n->in_idx_host) n->in_idx_host = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
        if (!n->in_mask_host) n->in_mask_host = (uint8_t *)malloc((size_t)N);
        if (!n->in_idx_host || !n->in_mask_host) return set_err(n->ctx, BGHIP_ENOMEM, "upload_input: host allocation");
        for (int i = 0; i < N; i++) {
            n->in_idx_host[i] = st && st->delayIdx ? st->delayIdx[i] : 0;
            n->in_mask_host[i] = mask[i];
        }
        CHK(h2d(n, dn.in_mask, mask.data(), N)); CHK(h2d(n, dn.in_nisi, nisi.data(), N));
        CHK(h2d(n, dn.in_psr, psr.data(), N)); CHK(h2d(n, dn.in_r, r.data(), N)); CHK(h2d(n, dn.in_u, u.data(), N));
        CHK(h2d(n, dn.in_last, last.data(), N)); CHK(h2d(n, dn.in_ring, ring.data(), N));
        HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
        dn.in_kind = BG_INPUT_POISSON;
        return BGHIP_OK;
    }
    if (in->kind == BGHIP_INPUT_REGULAR) {
        if (!in->values) return set_err(n->ctx, BGHIP_EINVAL, "upload_input: regular input needs values");
        // Inputs/SInputRegular.cpp:52-55
        const int dur = (int)((double)(in->duration / dt) + 0.5);
        const int itv = (int)((double)(in->interval / dt) + 0.5);
        if (dur < 0 || itv < 0 || dur + itv <= 0)
            return set_err(n->ctx, BGHIP_EINVAL, "upload_input: empty regular-input cycle");
        const int cycle = dur + itv;
        std::vector<int32_t> shift(N, 0);
        if (in->shift) std::copy(in->shift, in->shift + N, shift.begin());
        const int c_now = st ? st->steps_in_cycle : 0;
        if (c_now < 0 || c_now >= cycle) return set_err(n->ctx, BGHIP_EINVAL, "upload_input: steps_in_cycle out of range");
        dn.in_dur = dur;
        dn.in_cycle = cycle;
        dn.in_cyc_off = ((uint64_t)c_now + (uint64_t)cycle - n->step % (uint64_t)cycle) % (uint64_t)cycle;
        CHK(h2d(n, dn.in_values, in->values, N)); CHK(h2d(n, dn.in_shift, shift.data(), N));
        HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
        dn.in_kind = BG_INPUT_REGULAR;
        return BGHIP_OK;
    }
    return set_err(n->ctx, BGHIP_EINVAL, "upload_input: unknown input kind %d", in->kind);
}

extern "C" int bghip_download_input(bghip_net *n, bghip_input_state *st)
{
    if (!n || !st) return BGHIP_EINVAL;
    net_begin(n);
    DevNet &dn = n->dn;
    const int N = dn.N;
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    if (dn.in_kind == BG_INPUT_REGULAR)
        st->steps_in_cycle = (int32_t)((n->step + dn.in_cyc_off) % (uint64_t)dn.in_cycle);
    if (dn.in_kind != BG_INPUT_POISSON) return bghip_sync(n);
    std::vector<uint32_t> ring(N);
    CHK(d2h(n, st->nISIs, dn.in_nisi, N)); CHK(d2h(n, st->psr, dn.in_psr, N)); CHK(d2h(n, st->r, dn.in_r, N));
    CHK(d2h(n, st->u, dn.in_u, N)); CHK(d2h(n, st->lastSpike, dn.in_last, N));
    CHK(d2h(n, ring.data(), dn.in_ring, N));
    CHK(bghip_sync(n));
    const uint32_t phase = n->in_phase & 31u;
    const int di = (int)((n->step - (uint64_t)phase) & 31u);
    for (int i = 0; i < N; i++) {
        const bool on = n->in_mask_host[i] != 0;
        if (st->delayQueue)
            st->delayQueue[i] = on && phase ? (ring[i] >> phase) | (ring[i] << (32u - phase)) : ring[i];
        if (st->delayIdx) st->delayIdx[i] = on ? di : n->in_idx_host[i];
    }
    return BGHIP_OK;
}

// ------------------------------------------------------------------------------------------------
// Checkpoints: a versioned binary image of one network. The arena is copied as raw bytes (its
// layout is a function of the description alone, so the loading library re-derives every device
// pointer); the host-side state that is not in the arena follows as explicit little-endian fields.
// Philox noise and Poisson draws are functions of (seed, neuron, step), so the step counter is the
// whole random-number state.
static const char kCkptMagic[8] = {'B', 'G', 'H', 'I', 'P', 'N', 'E', 'T'};
static const uint32_t kCkptVersion = 7;

struct CkptHost {   // fixed-size record of the host-side state
    uint64_t step, last_t0, total, hist_step0, in_cyc_off;
    int64_t n_burst, n_shist;
    int32_t kmax, sel, min_delay_uploaded, ds_typed, in_kind, in_dur, in_cycle, rec_epochs, rec_epoch;
    uint32_t in_phase;
    uint8_t pending, has_layout, has_growth, rec_on, rec_growth, hist_range_known, has_in_host, pad;
    bghip_growth growth;
    float in_W, in_lambda, in_decay[2], in_U[2], in_D[2], in_F[2];
    int32_t in_delay[2];
    int32_t impl;            // implicit rings (the fire words and cohort states are in the arena)
    uint32_t coh_alloc;
    uint64_t young_until;
    int32_t stdp_typed;      // the per-type STDP table
    uint8_t tp_set[4];
    StdpParams tp[4];
};

template <typename T>
static bool put(FILE *f, const T *p, size_t n) { return n == 0 || fwrite(p, sizeof(T), n, f) == n; }
template <typename T>
static bool get(FILE *f, T *p, size_t n) { return n == 0 || fread(p, sizeof(T), n, f) == n; }

extern "C" int bghip_net_save(bghip_net *n, const char *path)
{
    if (!n || !path) return BGHIP_EINVAL;
    net_begin(n);
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    CHK(bghip_sync(n));
    flush_pending(n);   // a pending block's spikes go into the rings first
    CHK(coh_collect(n));
    const DevNet &dn = n->dn;
    CkptHost h;
    memset(&h, 0, sizeof h);
    h.impl = dn.impl; h.coh_alloc = n->coh_alloc; h.young_until = dn.young_until;
    h.step = n->step; h.last_t0 = n->last_t0; h.total = net_total(n); h.hist_step0 = n->hist_step0;
    h.in_cyc_off = dn.in_cyc_off; h.n_burst = n->n_burst; h.n_shist = n->n_shist;
    h.kmax = n->kmax; h.sel = n->sel; h.min_delay_uploaded = n->min_delay_uploaded; h.ds_typed = dn.ds_typed;
    h.stdp_typed = dn.stdp_typed;
    for (int t = 0; t < 4; t++) { h.tp_set[t] = n->tp_set[t] ? 1 : 0; h.tp[t] = dn.tp[t]; }
    h.in_kind = dn.in_kind; h.in_dur = dn.in_dur; h.in_cycle = dn.in_cycle; h.rec_epochs = n->rec_epochs;
    h.rec_epoch = n->rec_epoch; h.in_phase = n->in_phase; h.pending = n->pending; h.has_layout = n->has_layout;
    h.has_growth = n->has_growth; h.rec_on = n->rec_on; h.rec_growth = n->rec_growth;
    h.hist_range_known = n->hist_range_known; h.has_in_host = n->in_idx_host != nullptr;
    h.growth = n->growth; h.in_W = dn.in_W; h.in_lambda = dn.in_lambda;
    for (int k = 0; k < 2; k++) {
        h.in_decay[k] = dn.in_decay[k]; h.in_U[k] = dn.in_U[k]; h.in_D[k] = dn.in_D[k]; h.in_F[k] = dn.in_F[k];
        h.in_delay[k] = dn.in_delay[k];
    }
    std::vector<char> arena(n->arena_bytes);
    HIP_TRY(n->ctx, hipMemcpyAsync(arena.data(), n->arena, n->arena_bytes, hipMemcpyDeviceToHost, n->stream));
    const size_t N = dn.N, rows = (size_t)n->rec_epochs + 1;
    std::vector<uint32_t> burst(n->rec_on ? (size_t)n->n_burst : 0), shist(n->rec_on ? (size_t)n->n_shist : 0);
    std::vector<float> rh(n->rec_on && n->rec_growth ? rows * N : 0), qh(rh.size());
    if (n->rec_on) {
        CHK(d2h(n, burst.data(), n->d_burst, burst.size()));
        CHK(d2h(n, shist.data(), n->d_shist, shist.size()));
        CHK(d2h(n, rh.data(), n->d_rates_hist, rh.size()));
        CHK(d2h(n, qh.data(), n->d_radii_hist, qh.size()));
    }
    HIP_TRY(n->ctx, hipStreamSynchronize(n->stream));
    FILE *f = fopen(path, "wb");
    if (!f) return set_err(n->ctx, BGHIP_EINVAL, "net_save: cannot open %s", path);
    const uint32_t sizes[3] = {(uint32_t)sizeof(bghip_net_desc), (uint32_t)sizeof(CkptHost), 0u};
    const uint64_t ab = n->arena_bytes;
    bool ok = put(f, kCkptMagic, 8) && put(f, &kCkptVersion, 1) && put(f, sizes, 3) && put(f, &n->desc, 1) &&
              put(f, &h, 1) && put(f, &ab, 1) && put(f, arena.data(), arena.size());
    if (ok && h.has_in_host) ok = put(f, n->in_idx_host, N) && put(f, n->in_mask_host, N);
    if (ok && n->rec_on)
        ok = put(f, burst.data(), burst.size()) && put(f, shist.data(), shist.size()) && put(f, rh.data(), rh.size()) &&
             put(f, qh.data(), qh.size());
    ok = (fclose(f) == 0) && ok;
    return ok ? BGHIP_OK : set_err(n->ctx, BGHIP_EINVAL, "net_save: write to %s failed", path);
}

extern "C" int bghip_net_load(bghip_ctx *ctx, const char *path, bghip_net **out)
{
    if (!ctx || !path || !out) return BGHIP_EINVAL;
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) return set_err(ctx, BGHIP_EINVAL, "net_load: cannot open %s", path);
    char magic[8];
    uint32_t version = 0, sizes[3] = {0, 0, 0};
    bghip_net_desc desc;
    CkptHost h;
    uint64_t ab = 0;
    bool ok = get(f, magic, 8) && memcmp(magic, kCkptMagic, 8) == 0 && get(f, &version, 1) &&
              version == kCkptVersion && get(f, sizes, 3) && sizes[0] == sizeof(bghip_net_desc) &&
              sizes[1] == sizeof(CkptHost) && get(f, &desc, 1) && get(f, &h, 1) && get(f, &ab, 1);
    if (!ok) {
        fclose(f);
        return set_err(ctx, BGHIP_EINVAL, "net_load: %s is not a network checkpoint of this version", path);
    }
    bghip_net *n = nullptr;
    int rc = bghip_net_create(ctx, &desc, &n);
    if (rc) { fclose(f); return rc; }
    const size_t N = n->dn.N;
    std::vector<char> arena(ab);
    ok = ab == n->arena_bytes && get(f, arena.data(), arena.size());
    std::vector<int32_t> idx(h.has_in_host ? N : 0);
    std::vector<uint8_t> msk(h.has_in_host ? N : 0);
    if (ok && h.has_in_host) ok = get(f, idx.data(), N) && get(f, msk.data(), N);
    const size_t rows = (size_t)h.rec_epochs + 1;
    std::vector<uint32_t> burst(h.rec_on ? (size_t)h.n_burst : 0), shist(h.rec_on ? (size_t)h.n_shist : 0);
    std::vector<float> rh(h.rec_on && h.rec_growth ? rows * N : 0), qh(rh.size());
    if (ok && h.rec_on)
        ok = get(f, burst.data(), burst.size()) && get(f, shist.data(), shist.size()) && get(f, rh.data(), rh.size()) &&
             get(f, qh.data(), qh.size());
    ok = ok && fgetc(f) == EOF;
    fclose(f);
    if (!ok) {
        bghip_net_destroy(n);
        return set_err(ctx, BGHIP_EINVAL, "net_load: %s is truncated or does not match its description", path);
    }
    DevNet &dn = n->dn;
    if (hipMemcpy(n->arena, arena.data(), ab, hipMemcpyHostToDevice) != hipSuccess) {
        bghip_net_destroy(n);
        return set_err(ctx, BGHIP_EDEVICE, "net_load: upload failed");
    }
    n->step = h.step; n->last_t0 = h.last_t0; n->total = h.total; n->total_pending = false;
    dn.impl = h.impl; n->coh_alloc = h.coh_alloc; dn.coh_mask = h.coh_alloc; dn.young_until = h.young_until;
    n->hist_step0 = h.hist_step0; n->kmax = h.kmax; dn.look = h.kmax; n->sel = h.sel; n->min_delay_uploaded = h.min_delay_uploaded;
    n->pending = h.pending; n->has_layout = h.has_layout; n->has_growth = h.has_growth; n->growth = h.growth;
    n->hist_range_known = h.hist_range_known; n->in_phase = h.in_phase; n->out_valid = false;
    dn.ds_typed = h.ds_typed; dn.in_kind = h.in_kind; dn.in_dur = h.in_dur; dn.in_cycle = h.in_cycle;
    dn.stdp_typed = h.stdp_typed;
    for (int t = 0; t < 4; t++) { n->tp_set[t] = h.tp_set[t] != 0; dn.tp[t] = h.tp[t]; }
    dn.in_cyc_off = h.in_cyc_off; dn.in_W = h.in_W; dn.in_lambda = h.in_lambda;
    for (int k = 0; k < 2; k++) {
        dn.in_decay[k] = h.in_decay[k]; dn.in_U[k] = h.in_U[k]; dn.in_D[k] = h.in_D[k]; dn.in_F[k] = h.in_F[k];
        dn.in_delay[k] = h.in_delay[k];
    }
    if (h.has_in_host) {
        n->in_idx_host = (int32_t *)malloc(sizeof(int32_t) * N);
        n->in_mask_host = (uint8_t *)malloc(N);
        if (!n->in_idx_host || !n->in_mask_host) {
            bghip_net_destroy(n);
            return set_err(ctx, BGHIP_ENOMEM, "net_load: host allocation");
        }
        memcpy(n->in_idx_host, idx.data(), sizeof(int32_t) * N);
        memcpy(n->in_mask_host, msk.data(), N);
    }
    if (h.rec_on) {
        rc = bghip_recorder_enable(n, h.n_burst, h.n_shist, h.rec_epochs, h.rec_growth, nullptr, nullptr);
        if (rc) { bghip_net_destroy(n); return rc; }
        n->rec_epoch = h.rec_epoch;
        if (h2d(n, n->d_burst, burst.data(), burst.size()) || h2d(n, n->d_shist, shist.data(), shist.size()) ||
            h2d(n, n->d_rates_hist, rh.data(), rh.size()) || h2d(n, n->d_radii_hist, qh.data(), qh.size()) ||
            hipStreamSynchronize(n->stream) != hipSuccess) {
            bghip_net_destroy(n);
            return set_err(ctx, BGHIP_EDEVICE, "net_load: recorder upload failed");
        }
    }
    if (n->has_layout) {   // the layout's neuron types, for synapses added later
        n->ntype_host.resize((size_t)dn.N);
        if (d2h(n, n->ntype_host.data(), dn.ntype, (size_t)dn.N) || hipStreamSynchronize(n->stream) != hipSuccess) {
            bghip_net_destroy(n);
            return set_err(ctx, BGHIP_EDEVICE, "net_load: layout download failed");
        }
    }
    *out = n;
    return BGHIP_OK;
}

// ------------------------------------------------------------------------------------------------
// Batched form.
// How a network advances: neurons only (no synapses), fused temporal blocks, or per step.
enum { BG_MODE_QUIET = 0, BG_MODE_FUSED = 1, BG_MODE_STEP = 2 };
static int net_mode(const bghip_net *x)
{
    if (net_total(const_cast<bghip_net *>(x)) == 0) return BG_MODE_QUIET;
    const int sk = x->desc.synapse_model;
    return (x->kmax <= 1 || sk == BGHIP_SYN_STDP || sk == BGHIP_SYN_DYNSTDP) ? BG_MODE_STEP : BG_MODE_FUSED;
}

struct BatchGroup {
    int impl = 0;   // fused groups: implicit-ring members
    int nk, sk, w32;
    int mode;
    std::vector<int> members;      // indices into bghip_batch::nets
    BatchNet *d_nets = nullptr;    // device table (nn entries)
    uint32_t *d_base = nullptr;    // neuron prefix (nn + 1)
    uint16_t *d_nof = nullptr;     // member of each neuron / destination (total)
    uint2 *d_chunks[3] = {};       // per-step STDP: runs of 1, 2 and 4 destinations per wavefront
    uint32_t nchunks[3] = {};
    uint2 *d_nwaves = nullptr;     // thread-per-neuron kernels: {member, first neuron} per wavefront
    uint32_t nnwaves = 0;
    uint2 *d_quads = nullptr;      // implicit-ring fused kernels: {member, first position} per workgroup
    uint32_t nquads = 0;
    BatchShape shape = {};         // launch shape in use
    bool tuned = false;            // shape picked by the autotuner (or set by the caller)
    float step_us = 0.0f;          // measured time per step of that shape
    hipGraphExec_t g_exec = nullptr;   // per-step groups: BG_GRAPH_STEPS captured steps of this shape
    BatchShape g_shape = {};
    uint64_t *d_sbase = nullptr;   // slot prefix (nn + 1)
    uint64_t *d_tprev = nullptr;   // flush parameters (nn)
    int *d_sel = nullptr;
    EvLists ev = {};               // per-step STDP: sharded event lists of the group
    uint32_t *d_ev = nullptr;      // their counters, offsets and destination list in one allocation
    uint2 *d_evs = nullptr;        // their synapse list (sum of the members' slots)
    uint32_t total = 0;
    int kmax = 1;
    // pinned staging of the per-call tables, so an advance call does not wait for the device;
    // `staged` marks the copies of the previous call, which must finish before the next rewrite
    BatchNet *h_nets = nullptr;
    uint64_t *h_tprev = nullptr;
    int *h_sel = nullptr;
    hipEvent_t staged = nullptr;
    bool staged_pending = false;
};

struct bghip_batch {
    bghip_ctx *ctx;
    hipStream_t stream;
    hipEvent_t done;
    hipEvent_t tev0, tev1;         // autotuner trial timing
    uint32_t tune_steps = 0;       // autotuner: timed steps per trial (0: off)
    std::vector<bghip_net *> nets;
    std::vector<BatchGroup> groups;
    std::vector<int> modes;        // net_mode of each member when the groups were formed
    std::vector<int> kmaxs;        // and its block length (a group's is the minimum of its members')
    std::vector<int> impls;        // and whether it ran with implicit rings
    bool stale = false;            // a member was destroyed: the groups must be formed again
};

static void free_groups(bghip_batch *b)
{
    for (auto &g : b->groups) {
        if (g.g_exec) (void)hipGraphExecDestroy(g.g_exec);
        (void)hipFree(g.d_nets);
        (void)hipFree(g.d_base);
        (void)hipFree(g.d_nof);
        for (uint2 *c : g.d_chunks) (void)hipFree(c);
        (void)hipFree(g.d_nwaves);
        (void)hipFree(g.d_quads);
        (void)hipFree(g.d_sbase);
        (void)hipFree(g.d_tprev);
        (void)hipFree(g.d_sel);
        (void)hipFree(g.d_ev);
        (void)hipFree(g.d_evs);
        (void)hipHostFree(g.h_nets);
        (void)hipHostFree(g.h_tprev);
        (void)hipHostFree(g.h_sel);
        if (g.staged) (void)hipEventDestroy(g.staged);
    }
    b->groups.clear();
}

// A member being destroyed: the batch finishes its queued work, drops the network and its group
// tables (which hold a copy of the network's device descriptor), and forms its groups again at
// the next call.
static void batch_forget_net(bghip_batch *b, const bghip_net *n)
{
    auto it = std::find(b->nets.begin(), b->nets.end(), n);
    if (it == b->nets.end()) return;
    (void)hipStreamSynchronize(b->stream);
    b->nets.erase(it);
    free_groups(b);
    b->stale = true;
}

extern "C" int bghip_batch_create(bghip_ctx *ctx, bghip_batch **out)
{
    if (!ctx || !out) return BGHIP_EINVAL;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    bghip_batch *b = new (std::nothrow) bghip_batch();
    if (!b) return BGHIP_ENOMEM;
    b->ctx = ctx;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->tev0) != hipSuccess || hipEventCreate(&b->tev1) != hipSuccess) {
        delete b;
        return set_err(ctx, BGHIP_EDEVICE, "batch stream/event creation failed");
    }
    ctx->batches.push_back(b);
    *out = b;
    return BGHIP_OK;
}

extern "C" void bghip_batch_destroy(bghip_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->stream);
    auto &bs = b->ctx->batches;
    bs.erase(std::remove(bs.begin(), bs.end(), b), bs.end());
    free_groups(b);
    (void)hipEventDestroy(b->tev0);
    (void)hipEventDestroy(b->tev1);
    (void)hipEventDestroy(b->done);
    (void)hipStreamDestroy(b->stream);
    delete b;
}

// The untuned launch shape: 256-thread blocks, four destinations per streaming wavefront, event
// grids growing with the batch (up to 4 and 8 blocks per shard, one per 64 destinations a shard
// holds on average).
static BatchShape default_shape(uint32_t total)
{
    const uint64_t per = 64 * BG_EV_SHARDS_BATCH;
    const int g = (int)std::max<uint64_t>(1, ((uint64_t)total + per - 1) / per);
    return BatchShape{BG_BLOCK, BG_STREAM_DESTS, BG_BLOCK, std::min(4, g), std::min(8, g)};
}

static int size_class(uint32_t total)
{
    int c = 0;
    while (total > 1u) { total >>= 1; c++; }
    return c;
}

static const TuneEntry *find_tuned(const bghip_ctx *ctx, const BatchGroup &g)
{
    for (const TuneEntry &e : ctx->tuned)
        if (e.mode == g.mode && e.nk == g.nk && e.sk == g.sk && e.w32 == g.w32 && e.size_class == size_class(g.total))
            return &e;
    return nullptr;
}

static bool shape_valid(const BatchShape &s)
{
    const auto blk_ok = [](int v) { return v == 64 || v == 128 || v == 256; };
    return blk_ok(s.blk) && blk_ok(s.blk_stream) && (s.dpw == 1 || s.dpw == 2 || s.dpw == 4) && s.g_syn >= 1 &&
           s.g_syn <= 64 && s.g_sum >= 1 && s.g_sum <= 64;
}

// Groups the members by (neuron model, synapse model, ring words, mode); each group advances with
// shared launches and its own device tables.
static int regroup(bghip_batch *b)
{
    HIP_TRY(b->ctx, hipStreamSynchronize(b->stream));
    free_groups(b);
    b->stale = false;
    const int n = (int)b->nets.size();
    b->modes.assign(n, 0);
    b->kmaxs.assign(n, 0);
    b->impls.assign(n, 0);
    for (int i = 0; i < n; i++) {
        bghip_net *x = b->nets[i];
        const int mode = net_mode(x);
        b->modes[i] = mode;
        b->kmaxs[i] = x->kmax;
        b->impls[i] = x->dn.impl;
        const int sk = mode == BG_MODE_QUIET ? -1 : x->desc.synapse_model;
        const int w32 = mode == BG_MODE_QUIET ? 0 : x->dn.W32;
        const int impl = mode == BG_MODE_FUSED ? x->dn.impl : 0;
        BatchGroup *grp = nullptr;
        for (auto &g : b->groups)
            if (g.nk == x->desc.neuron_model && g.sk == sk && g.w32 == w32 && g.mode == mode && g.impl == impl)
                grp = &g;
        if (!grp) {
            BatchGroup g;
            g.impl = impl;
            g.nk = x->desc.neuron_model;
            g.sk = sk;
            g.w32 = w32;
            g.mode = mode;
            g.kmax = BG_KMAX;
            b->groups.push_back(g);
            grp = &b->groups.back();
        }
        grp->members.push_back(i);
        grp->kmax = std::min(grp->kmax, x->kmax);
    }
    for (auto &g : b->groups) {
        const int nn = (int)g.members.size();
        std::vector<uint32_t> base(nn + 1, 0);
        std::vector<uint64_t> sbase(nn + 1, 0);
        for (int k = 0; k < nn; k++) {
            const DevNet &dn = b->nets[g.members[k]]->dn;
            base[k + 1] = base[k] + (uint32_t)dn.N;
            sbase[k + 1] = sbase[k] + dn.S;
        }
        g.total = base[nn];
        if (nn > 65535) return set_err(b->ctx, BGHIP_EUNSUPPORTED, "more than 65535 networks of one kind in a batch");
        std::vector<uint16_t> nof(std::max<uint32_t>(g.total, 1), 0);
        for (int k = 0; k < nn; k++) std::fill(nof.begin() + base[k], nof.begin() + base[k + 1], (uint16_t)k);
        HIP_TRY(b->ctx, hipMalloc((void **)&g.d_nof, sizeof(uint16_t) * nof.size()));
        HIP_TRY(b->ctx, hipMemcpy(g.d_nof, nof.data(), sizeof(uint16_t) * nof.size(), hipMemcpyHostToDevice));
        std::vector<uint2> nwaves;
        for (int k = 0; k < nn; k++)
            for (uint32_t first = 0; first < base[k + 1] - base[k]; first += BG_WAVE) nwaves.push_back(make_uint2(k, first));
        g.nnwaves = (uint32_t)nwaves.size();
        HIP_TRY(b->ctx, hipMalloc((void **)&g.d_nwaves, sizeof(uint2) * std::max<size_t>(nwaves.size(), 1)));
        HIP_TRY(b->ctx, hipMemcpy(g.d_nwaves, nwaves.data(), sizeof(uint2) * nwaves.size(), hipMemcpyHostToDevice));
        std::vector<uint2> quads;
        if (g.impl)
            for (int k = 0; k < nn; k++)
                for (uint32_t first = 0; first < base[k + 1] - base[k]; first += BG_BLOCK / BG_WAVE)
                    quads.push_back(make_uint2(k, first));
        g.nquads = (uint32_t)quads.size();
        HIP_TRY(b->ctx, hipMalloc((void **)&g.d_quads, sizeof(uint2) * std::max<size_t>(quads.size(), 1)));
        if (!quads.empty())
            HIP_TRY(b->ctx, hipMemcpy(g.d_quads, quads.data(), sizeof(uint2) * quads.size(), hipMemcpyHostToDevice));
        HIP_TRY(b->ctx, hipMalloc((void **)&g.d_nets, sizeof(BatchNet) * nn));
        HIP_TRY(b->ctx, hipMalloc((void **)&g.d_base, sizeof(uint32_t) * (nn + 1)));
        HIP_TRY(b->ctx, hipMalloc((void **)&g.d_sbase, sizeof(uint64_t) * (nn + 1)));
        HIP_TRY(b->ctx, hipMalloc((void **)&g.d_tprev, sizeof(uint64_t) * nn));
        HIP_TRY(b->ctx, hipMalloc((void **)&g.d_sel, sizeof(int) * nn));
        HIP_TRY(b->ctx, hipHostMalloc((void **)&g.h_nets, sizeof(BatchNet) * nn, hipHostMallocDefault));
        HIP_TRY(b->ctx, hipHostMalloc((void **)&g.h_tprev, sizeof(uint64_t) * nn, hipHostMallocDefault));
        HIP_TRY(b->ctx, hipHostMalloc((void **)&g.h_sel, sizeof(int) * nn, hipHostMallocDefault));
        HIP_TRY(b->ctx, hipEventCreateWithFlags(&g.staged, hipEventDisableTiming));
        if (g.mode == BG_MODE_STEP && (g.sk == BGHIP_SYN_STDP || g.sk == BGHIP_SYN_DYNSTDP)) {
            const uint64_t NS = BG_EV_SHARDS_BATCH;
            std::vector<EvPart> parts(nn);
            for (int k = 0; k < nn; k++) {
                const DevNet &dn = b->nets[g.members[k]]->dn;
                parts[k] = {(uint64_t)dn.N, (uint64_t)dn.M, 0, 1};
            }
            ev_batch_parts(parts, NS);
            std::vector<uint32_t> doff, soff;
            if (!ev_offsets(parts, NS, doff, soff))
                return set_err(b->ctx, BGHIP_EUNSUPPORTED, "batch too large for the event lists");
            const size_t nc = 2 * NS * BG_EV_PAD, no = 2 * (NS + 1);
            HIP_TRY(b->ctx, hipMalloc((void **)&g.d_ev, sizeof(uint32_t) * (nc + no + (size_t)g.total + 1)));
            HIP_TRY(b->ctx, hipMalloc((void **)&g.d_evs, sizeof(uint2) * (size_t)std::max<uint32_t>(soff[NS], 1)));
            doff.insert(doff.end(), soff.begin(), soff.end());
            HIP_TRY(b->ctx, hipMemset(g.d_ev, 0, sizeof(uint32_t) * nc));
            HIP_TRY(b->ctx, hipMemcpy(g.d_ev + nc, doff.data(), sizeof(uint32_t) * no, hipMemcpyHostToDevice));
            g.ev.count = g.d_ev;
            g.ev.dest_off = g.d_ev + nc;
            g.ev.syn_off = g.d_ev + nc + NS + 1;
            g.ev.dest = g.d_ev + nc + no;
            g.ev.syn = g.d_evs;
            // runs of D = 1, 2, 4 destinations; a run's shard is that of its BG_STREAM_DESTS-run
            // (ev_shard of chunk d / BG_STREAM_DESTS), the assignment the capacities were sized for
            for (int v = 0; v < 3; v++) {
                const uint32_t D = 1u << v;
                std::vector<uint2> chunks;
                for (int k = 0; k < nn; k++)
                    for (uint32_t d = 0; d < base[k + 1] - base[k]; d += D) {
                        const uint32_t shard = ev_shard(parts[k], d / BG_STREAM_DESTS, NS);
                        chunks.push_back(make_uint2((uint32_t)k | (shard << 16), d));
                    }
                g.nchunks[v] = (uint32_t)chunks.size();
                HIP_TRY(b->ctx, hipMalloc((void **)&g.d_chunks[v], sizeof(uint2) * std::max<size_t>(chunks.size(), 1)));
                HIP_TRY(b->ctx, hipMemcpy(g.d_chunks[v], chunks.data(), sizeof(uint2) * chunks.size(),
                                          hipMemcpyHostToDevice));
            }
        }
        g.shape = default_shape(g.total);
        if (const TuneEntry *te = find_tuned(b->ctx, g)) {
            g.shape = te->shape;
            g.step_us = te->step_us;
            g.tuned = true;
        }
        HIP_TRY(b->ctx, hipMemcpy(g.d_base, base.data(), sizeof(uint32_t) * (nn + 1), hipMemcpyHostToDevice));
        HIP_TRY(b->ctx, hipMemcpy(g.d_sbase, sbase.data(), sizeof(uint64_t) * (nn + 1), hipMemcpyHostToDevice));
    }
    return BGHIP_OK;
}

// True if a member changed mode or block length since the groups were formed.
static bool members_changed(const bghip_batch *b)
{
    if (b->stale || b->modes.size() != b->nets.size()) return true;
    for (size_t i = 0; i < b->nets.size(); i++)
        if (net_mode(b->nets[i]) != b->modes[i] || b->nets[i]->kmax != b->kmaxs[i] ||
            b->nets[i]->dn.impl != b->impls[i])
            return true;
    return false;
}

static BatchMap batch_map(const BatchGroup &g, const BatchShape &sh)
{
    const int v = sh.dpw == 1 ? 0 : (sh.dpw == 2 ? 1 : 2);
    return BatchMap{g.d_nets, g.d_base,      g.d_nof,    g.total, (int)g.members.size(), g.d_chunks[v],
                    g.nchunks[v], g.d_nwaves, g.nnwaves, g.d_quads, g.nquads};
}

// ---- batch autotuner -----------------------------------------------------------------------
// One axis of the search: a field of the shape and the values tried for it.
struct ShapeAxis {
    int BatchShape::*field;
    std::vector<int> values;
};

static std::vector<ShapeAxis> quiet_space() { return {{&BatchShape::blk, {256, 128, 64}}}; }
static std::vector<ShapeAxis> fused_space() { return {{&BatchShape::blk, {256, 128, 64}}}; }
// implicit-ring fused groups run one workgroup of BG_BLOCK threads per four destinations: the one
// shape is still timed, so the group reports its time per step
static std::vector<ShapeAxis> fused_impl_space() { return {{&BatchShape::blk, {BG_BLOCK}}}; }
static std::vector<ShapeAxis> step_space(const BatchGroup &g)
{
    if (g.sk != BGHIP_SYN_STDP && g.sk != BGHIP_SYN_DYNSTDP) return {{&BatchShape::blk, {256, 128, 64}}};
    return {{&BatchShape::dpw, {4, 2, 1}},
            {&BatchShape::blk_stream, {256, 128, 64}},
            {&BatchShape::g_syn, {1, 2, 4, 8, 16}},
            {&BatchShape::g_sum, {1, 2, 4, 8, 16, 32}},
            {&BatchShape::blk, {256, 128, 64}}};
}
static int space_size(const std::vector<ShapeAxis> &sp)
{
    int n = 0;
    for (const ShapeAxis &a : sp) n += (int)a.values.size();
    return n;
}

// Runs `work` (which enqueues `steps` steps) between two events on the batch stream and returns
// the time per step in microseconds.
template <class F>
static int time_steps(bghip_batch *b, float &us, uint64_t steps, F &&work)
{
    HIP_TRY(b->ctx, hipEventRecord(b->tev0, b->stream));
    work();
    HIP_TRY(b->ctx, hipGetLastError());
    HIP_TRY(b->ctx, hipEventRecord(b->tev1, b->stream));
    HIP_TRY(b->ctx, hipEventSynchronize(b->tev1));
    float ms = 0.0f;
    HIP_TRY(b->ctx, hipEventElapsedTime(&ms, b->tev0, b->tev1));
    us = ms * 1e3f / (float)std::max<uint64_t>(1, steps);
    return BGHIP_OK;
}

// Coordinate descent over the axes, starting from the group's current shape: every candidate value
// of an axis is tried with the other fields fixed at the best found so far, and kept if faster.
// Trials run real steps of the member networks; every shape gives bit-identical results, so the
// search changes only the time the steps take. The result is remembered per device for groups of
// the same kind and size class.
template <class Trial>
static int tune_group(bghip_batch *b, BatchGroup &g, const std::vector<ShapeAxis> &space, Trial &&trial)
{
    BatchShape cur = g.shape;
    float best = 0.0f;
    CHK(trial(cur, best));
    for (const ShapeAxis &a : space)
        for (int v : a.values) {
            if (cur.*(a.field) == v) continue;
            BatchShape c = cur;
            c.*(a.field) = v;
            float us = 0.0f;
            CHK(trial(c, us));
            if (us < best) {
                best = us;
                cur = c;
            }
        }
    g.shape = cur;
    g.step_us = best;
    g.tuned = true;
    const TuneEntry e = {g.mode, g.nk, g.sk, g.w32, size_class(g.total), cur, best};
    auto &tab = b->ctx->tuned;
    auto it = std::find_if(tab.begin(), tab.end(), [&](const TuneEntry &x) {
        return x.mode == e.mode && x.nk == e.nk && x.sk == e.sk && x.w32 == e.w32 && x.size_class == e.size_class;
    });
    if (it != tab.end()) *it = e;
    else tab.push_back(e);
    return BGHIP_OK;
}

// Captures BG_GRAPH_STEPS steps of a per-step group (offsets 0..S-1 from each member's table step)
// and the table bump as one executable graph, replayed for every S steps of later calls.
static int capture_batch_graph(bghip_batch *b, BatchGroup &g)
{
    const BatchMap bm = batch_map(g, g.shape);
    hipGraph_t gr = nullptr;
    HIP_TRY(b->ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t i = 0; i < BG_GRAPH_STEPS; i++) launch_step_batch(bm, b->stream, g.shape, g.nk, g.sk, g.w32, i, g.ev);
    launch_batch_bump(b->stream, g.d_nets, (int)g.members.size(), BG_GRAPH_STEPS);
    const hipError_t e = hipStreamEndCapture(b->stream, &gr);
    if (e != hipSuccess) return set_err(b->ctx, BGHIP_EDEVICE, "batch graph capture failed: %s", hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&g.g_exec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e2 != hipSuccess) {
        g.g_exec = nullptr;
        return set_err(b->ctx, BGHIP_EDEVICE, "batch graph instantiation failed: %s", hipGetErrorString(e2));
    }
    g.g_shape = g.shape;
    return BGHIP_OK;
}

// Steps s .. n_steps-1 of this call for a per-step group: whole runs of BG_GRAPH_STEPS replay the
// captured graph (after one bump that moves the table's steps past the s steps already run), the
// rest is launched directly.
static int run_step_group(bghip_batch *b, BatchGroup &g, uint64_t s, uint64_t n_steps)
{
    const BatchMap bm = batch_map(g, g.shape);
    uint64_t off = s;   // offset of step s from the steps in the device table
    if (n_steps - s >= BG_GRAPH_STEPS && b->ctx->graphs) {
        if (g.g_exec && memcmp(&g.g_shape, &g.shape, sizeof(BatchShape)) != 0) {
            (void)hipGraphExecDestroy(g.g_exec);
            g.g_exec = nullptr;
        }
        if (!g.g_exec) CHK(capture_batch_graph(b, g));
        if (off) launch_batch_bump(b->stream, g.d_nets, (int)g.members.size(), (uint32_t)off);
        off = 0;
        for (; s + BG_GRAPH_STEPS <= n_steps; s += BG_GRAPH_STEPS) HIP_TRY(b->ctx, hipGraphLaunch(g.g_exec, b->stream));
    }
    for (; s < n_steps; s++, off++) launch_step_batch(bm, b->stream, g.shape, g.nk, g.sk, g.w32, off, g.ev);
    return BGHIP_OK;
}

extern "C" int bghip_batch_set(bghip_batch *b, bghip_net *const *nets, int n)
{
    if (!b || n < 0 || (n > 0 && !nets)) return BGHIP_EINVAL;
    HIP_TRY(b->ctx, hipSetDevice(b->ctx->device));
    for (int i = 0; i < n; i++)
        if (!nets[i] || nets[i]->ctx->device != b->ctx->device)
            return set_err(b->ctx, BGHIP_EINVAL, "batch member %d invalid", i);
    b->nets.assign(nets, nets + n);
    return regroup(b);
}

extern "C" int bghip_batch_advance(bghip_batch *b, uint64_t n_steps)
{
    if (!b) return BGHIP_EINVAL;
    if (b->nets.empty() || n_steps == 0) return BGHIP_OK;
    HIP_TRY(b->ctx, hipSetDevice(b->ctx->device));
    // the batch stream waits for every member's own outstanding work
    for (bghip_net *x : b->nets) {
        net_begin(x);
        if (x->desc.synapse_model == BGHIP_SYN_STDP || x->desc.synapse_model == BGHIP_SYN_DYNSTDP)
            CHK(build_outgoing(x));
        HIP_TRY(b->ctx, hipEventRecord(x->ext_ev, x->stream));
        HIP_TRY(b->ctx, hipStreamWaitEvent(b->stream, x->ext_ev, 0));
    }
    // a growth update or an upload may have moved a member to another mode, and new synapses with
    // shorter delays may have shortened its blocks
    if (members_changed(b)) CHK(regroup(b));
    for (auto &g : b->groups) {
        const int nn = (int)g.members.size();
        if (g.staged_pending) {   // the previous call's copies out of the staging buffers
            HIP_TRY(b->ctx, hipEventSynchronize(g.staged));
            g.staged_pending = false;
        }
        BatchNet *tab = g.h_nets;
        for (int k = 0; k < nn; k++) {
            bghip_net *x = b->nets[g.members[k]];
            tab[k].n = x->dn;
            tab[k].t0 = x->step;
            tab[k].tprev0 = x->last_t0;
            tab[k].sel = x->sel;
            tab[k].pull = x->pending ? 1 : 0;
        }
        HIP_TRY(b->ctx, hipMemcpyAsync(g.d_nets, tab, sizeof(BatchNet) * nn, hipMemcpyHostToDevice, b->stream));
        const bool tune = b->tune_steps > 0 && !g.tuned;
        const uint64_t T = b->tune_steps;
        if (g.mode == BG_MODE_QUIET) {
            uint64_t done = 0;
            if (tune && n_steps >= 4 * 11 * T) {
                // trial: a launch of T steps (first launch of the shape), then a timed one of 10 T
                CHK(tune_group(b, g, quiet_space(), [&](const BatchShape &sh, float &us) {
                    const BatchMap m = batch_map(g, sh);
                    launch_neuron_only_batch(m, b->stream, sh, g.nk, done, (int)T);
                    done += T;
                    return time_steps(b, us, 10 * T, [&] {
                        launch_neuron_only_batch(m, b->stream, sh, g.nk, done, (int)(10 * T));
                        done += 10 * T;
                    });
                }));
            }
            const BatchMap bm = batch_map(g, g.shape);
            while (done < n_steps) {
                const int k = (int)std::min<uint64_t>(BG_QUIET_STEPS, n_steps - done);
                launch_neuron_only_batch(bm, b->stream, g.shape, g.nk, done, k);
                done += (uint64_t)k;
            }
            for (int k = 0; k < nn; k++) {
                bghip_net *x = b->nets[g.members[k]];
                x->step += n_steps;
                x->pending = false;
            }
        } else if (g.mode == BG_MODE_STEP) {
            uint64_t s = 0;
            const std::vector<ShapeAxis> space = step_space(g);
            if (tune && n_steps >= (uint64_t)(space_size(space) + 1) * (T + 1)) {
                // trial: one step (first launch of the shape), then T timed steps
                CHK(tune_group(b, g, space, [&](const BatchShape &sh, float &us) {
                    const BatchMap m = batch_map(g, sh);
                    launch_step_batch(m, b->stream, sh, g.nk, g.sk, g.w32, s++, g.ev);
                    return time_steps(b, us, T, [&] {
                        for (uint64_t i = 0; i < T; i++) launch_step_batch(m, b->stream, sh, g.nk, g.sk, g.w32, s++, g.ev);
                    });
                }));
            }
            CHK(run_step_group(b, g, s, n_steps));
            for (int k = 0; k < nn; k++) {
                bghip_net *x = b->nets[g.members[k]];
                x->step += n_steps;
                x->pending = false;
            }
        } else {
            const int Kfull = std::max(1, g.kmax);
            const uint64_t L = (n_steps + (uint64_t)Kfull - 1) / (uint64_t)Kfull;
            auto fused_launch = [&](const BatchMap &m, const BatchShape &sh, int K, int launch) {
                if (g.impl) launch_fused_impl_batch(m, b->stream, sh, g.nk, g.sk, g.kmax, K, Kfull, launch);
                else launch_fused_batch(m, b->stream, sh, g.nk, g.sk, g.w32, g.kmax, K, Kfull, launch);
            };
            uint64_t l = 0;
            const uint64_t TL = std::max<uint64_t>(4, (T + (uint64_t)Kfull - 1) / (uint64_t)Kfull);
            if (tune && L > 4 * (TL + 1)) {
                // trial: one launch (first of the shape), then TL timed launches of Kfull steps
                CHK(tune_group(b, g, g.impl ? fused_impl_space() : fused_space(), [&](const BatchShape &sh, float &us) {
                    const BatchMap m = batch_map(g, sh);
                    fused_launch(m, sh, Kfull, (int)l++);
                    return time_steps(b, us, TL * (uint64_t)Kfull, [&] {
                        for (uint64_t i = 0; i < TL; i++) fused_launch(m, sh, Kfull, (int)l++);
                    });
                }));
            }
            const BatchMap bm = batch_map(g, g.shape);
            for (; l < L; l++) {
                const int K = (int)std::min<uint64_t>((uint64_t)Kfull, n_steps - l * (uint64_t)Kfull);
                fused_launch(bm, g.shape, K, (int)l);
            }
            uint64_t *tprev = g.h_tprev;
            int *sel = g.h_sel;
            for (int k = 0; k < nn; k++) {
                bghip_net *x = b->nets[g.members[k]];
                tprev[k] = x->step + (L - 1) * (uint64_t)Kfull;
                sel[k] = x->sel ^ (int)(L & 1);
                x->last_t0 = tprev[k];
                x->sel = sel[k];
                x->step += n_steps;
                x->pending = false;
            }
            HIP_TRY(b->ctx, hipMemcpyAsync(g.d_tprev, tprev, sizeof(uint64_t) * nn, hipMemcpyHostToDevice, b->stream));
            HIP_TRY(b->ctx, hipMemcpyAsync(g.d_sel, sel, sizeof(int) * nn, hipMemcpyHostToDevice, b->stream));
            uint64_t slots = 0;
            for (int k = 0; k < nn; k++) slots += b->nets[g.members[k]]->dn.S;
            if (slots && !g.impl) {   // implicit rings have nothing to flush
                unsigned blocks = (unsigned)((slots + BG_BLOCK - 1) / BG_BLOCK);
                hipLaunchKernelGGL(k_flush_batch, dim3(blocks), dim3(BG_BLOCK), 0, b->stream, g.d_nets, g.d_sbase, nn,
                                   g.d_tprev, g.d_sel);
            }
        }
        HIP_TRY(b->ctx, hipEventRecord(g.staged, b->stream));
        g.staged_pending = true;
    }
    HIP_TRY(b->ctx, hipGetLastError());
    HIP_TRY(b->ctx, hipEventRecord(b->done, b->stream));
    for (bghip_net *x : b->nets) {
        HIP_TRY(b->ctx, hipStreamWaitEvent(x->stream, b->done, 0));
    }
    return BGHIP_OK;
}

extern "C" int bghip_batch_sync(bghip_batch *b)
{
    if (!b) return BGHIP_EINVAL;
    HIP_TRY(b->ctx, hipSetDevice(b->ctx->device));
    HIP_TRY(b->ctx, hipStreamSynchronize(b->stream));
    return BGHIP_OK;
}

extern "C" int bghip_batch_autotune(bghip_batch *b, uint32_t trial_steps)
{
    if (!b || trial_steps > 1000) return BGHIP_EINVAL;
    b->tune_steps = trial_steps;
    return BGHIP_OK;
}

extern "C" int bghip_batch_group_count(bghip_batch *b)
{
    if (!b) return BGHIP_EINVAL;
    // a growth update or an upload since the last call may have moved a member to another mode
    if (members_changed(b)) CHK(regroup(b));
    return (int)b->groups.size();
}

extern "C" int bghip_batch_get_shape(bghip_batch *b, int group, bghip_batch_shape *out)
{
    if (!b || !out || group < 0 || group >= (int)b->groups.size()) return BGHIP_EINVAL;
    const BatchGroup &g = b->groups[group];
    out->mode = g.mode;
    out->neuron_model = g.nk;
    out->synapse_model = g.sk;
    out->ring_words = g.w32;
    out->networks = (uint32_t)g.members.size();
    out->work_items = g.total;
    out->block = g.shape.blk;
    out->dests_per_wave = g.shape.dpw;
    out->block_stream = g.shape.blk_stream;
    out->syn_blocks_per_shard = g.shape.g_syn;
    out->sum_blocks_per_shard = g.shape.g_sum;
    out->tuned = g.tuned ? 1 : 0;
    out->step_us = g.step_us;
    return BGHIP_OK;
}

extern "C" int bghip_batch_set_shape(bghip_batch *b, int group, const bghip_batch_shape *in)
{
    if (!b || !in || group < 0 || group >= (int)b->groups.size()) return BGHIP_EINVAL;
    const BatchShape sh = {in->block, in->dests_per_wave, in->block_stream, in->syn_blocks_per_shard,
                           in->sum_blocks_per_shard};
    if (!shape_valid(sh)) return set_err(b->ctx, BGHIP_EINVAL, "invalid batch launch shape");
    BatchGroup &g = b->groups[group];
    g.shape = sh;
    g.tuned = true;
    g.step_us = 0.0f;
    return BGHIP_OK;
}

extern "C" int bghip_kernel_timing(bghip_net *n, int enable, double *avg_ms, int64_t *count)
{
    if (!n) return BGHIP_EINVAL;
    HIP_TRY(n->ctx, hipSetDevice(n->ctx->device));
    if (enable) {
        n->timed_launches = 0;
        HIP_TRY(n->ctx, hipEventRecord(n->ev0, n->stream));
        n->timing = true;
        return BGHIP_OK;
    }
    if (!n->timing) return BGHIP_EINVAL;
    HIP_TRY(n->ctx, hipEventRecord(n->ev1, n->stream));
    HIP_TRY(n->ctx, hipEventSynchronize(n->ev1));
    float ms = 0.0f;
    HIP_TRY(n->ctx, hipEventElapsedTime(&ms, n->ev0, n->ev1));
    n->timing = false;
    if (count) *count = n->timed_launches;
    if (avg_ms) *avg_ms = n->timed_launches ? (double)ms / (double)n->timed_launches : 0.0;
    return BGHIP_OK;
}
