// Fused block kernels with implicit delay rings, for Spiking/DS networks with 32-bit rings.
//
// The reference keeps a delay ring per synapse (AllSpikingSynapses::delayQueue,
// Synapses/AllSpikingSynapses.h:354-365): preSpikeHit sets bit delay ahead when the source fires
// (Synapses/AllSpikingSynapses.cpp:261-279) and advanceSynapse tests and clears the current bit
// (:242-254, 297-319). At a step boundary T the ring of a synapse therefore holds exactly the spikes
// its source fired at steps s in [T - total_delay, T) with s at or after the synapse's creation
// step (a new synapse starts with an empty ring). Here each neuron keeps the fire bits of its last
// 64 steps (rec[], written once per block by the neuron's own wavefront) and a synapse reads its
// deliveries of a block from its source's word: no ring is stored, read or written.
//
// AllDSSynapses::changePSR (Synapses/AllDSSynapses.cpp:227-247) advances r, u and lastSpike at
// each delivery from the interval since the previous one. All synapses of one source, one type
// (so one delay and one U, D, F) and one creation step start from the same createSynapse values
// and see the same deliveries, so their r, u, lastSpike are the same numbers at every step. The
// source's wavefront advances that state once per (cohort, target class) at the end of each block,
// and prepares (u, r) after the first delivery of the next block; a delivering synapse then adds
// ((W / decay) * u) * r exactly as changePSR does, with the same operands. A synapse that receives
// a second spike inside one block (a source refractory period shorter than the block) advances a
// private copy of that state for the later ones, with the same arithmetic.
//
// Downloads and every switch to the explicit kernels first write the per-synapse rings and DS
// state these representations stand for into the slot arrays (k_materialize), bit for bit the
// reference's.
#pragma once
#include "bghip_kernels.h"

// Occupancy target of the implicit-ring fused kernels (waves per SIMD): bounds their registers.
#ifndef BG_IMPL_WAVES
#define BG_IMPL_WAVES 6
#endif

__device__ __forceinline__ int meta_coh(uint32_t meta) { return (int)((meta >> BG_META_COH_SHIFT) & (BG_COH_MAX - 1)); }

// Bits k of a block starting at t0 whose delivery (at t0 + k, delay steps after the spike) comes
// from a spike at or after step E.
__device__ __forceinline__ uint32_t since_mask(uint64_t E, uint32_t delay, uint64_t t0)
{
    const uint64_t first = E + (uint64_t)delay;
    if (first <= t0) return 0xFFFFFFFFu;
    const uint64_t lim = first - t0;
    return lim >= 32 ? 0u : ~((1u << (uint32_t)lim) - 1u);
}

// The r, u, lastSpike update of AllDSSynapses::changePSR at a delivery at step t (the same
// expressions as psr_increment_regs).
__device__ __forceinline__ void ds_deliver(const DevNet &n, float &r, float &u, uint64_t &last, uint64_t t, float U,
                                           float D, float F)
{
    if (last != BG_NO_SPIKE) {
        const uint64_t gap = t - last;
        const float isi = ((gap >> 32) ? (float)gap : (float)(uint32_t)gap) * n.dt;
        r = 1.0f + (r * (1.0f - u) - 1.0f) * bg_expf(-isi / D);
        u = U + u * (1.0f - U) * bg_expf(-isi / F);
    }
    last = t;
}

struct ImplRank {
    uint32_t slot;
    uint32_t meta;
    float psr;
};

// Rank q of the row (clamped to its last rank, so every lane loads a valid slot).
template <bool DENSE>
__device__ __forceinline__ void impl_hot_load(const DevNet &n, uint32_t row, const uint32_t *imap, uint32_t q,
                                              ImplRank &h)
{
    h.slot = DENSE ? row + q : imap[q];
    h.meta = ld_off(n.meta, h.slot << 2);
    h.psr = ld_off(n.psr, h.slot << 2);
}

// What a delivering synapse needs, fetched for every rank ahead of its pass (at the usual delivery
// rates every cache line of these arrays is needed anyway): W / decay and, for DS, the prepared
// (u, r) of its cohort.
struct ImplCold {
    float Wd;
    float2 f;
};

template <int SK>
__device__ __forceinline__ void impl_cold_load(const DevNet &n, const float2 *fac, const ImplRank &h, ImplCold &c)
{
    c.Wd = ld_off(n.Wd, h.slot << 2);
    if (SK == SK_DS) {
        const uint32_t src = h.meta & BG_META_SRC_MASK;
        const uint32_t pair = ((h.meta >> (BG_META_COH_SHIFT - 1)) & ((BG_COH_MAX - 1) << 1)) |
                              ((h.meta >> BG_META_TYPE_SHIFT) & 1u);   // cohort * 2 + target class
        c.f = ld_off(fac, (__umul24(pair, (uint32_t)n.N) + src) << 3);
    } else {
        c.f = make_float2(0.0f, 0.0f);
    }
}

// One synapse through the K steps of a block: its deliveries are bits (64 - delay) .. of its
// source's fire word (bit k = delivery at t0 + k), masked to spikes after its creation in blocks
// that may reach back before it; then the PSR with the canonical per-step accumulation. tab holds
// each type's decay and 64 - total delay, cstep each cohort's creation step. A lane past the end
// of the row (valid false) adds exactly +0.0 to its partial sums, which leaves them bit-identical
// (they start at +0.0 and so are never -0.0).
template <int SK, int KMAX, bool TAIL>
__device__ __forceinline__ void impl_synapse(const DevNet &n, const TypeRegs &tr, const uint2 *tab,
                                             const uint64_t *cstep, const ImplRank &h, float &psr, uint64_t recent,
                                             const ImplCold &c, bool valid, uint64_t t0, int K, bool young,
                                             float (&acc)[KMAX], uint32_t &hit)
{
    if (!TAIL) valid = true;
    const int Kb = BG_K_EXACT(KMAX) ? KMAX : K;
    const uint32_t kmask = Kb >= 32 ? 0xFFFFFFFFu : ((1u << Kb) - 1u);
    const int type = (h.meta >> BG_META_TYPE_SHIFT) & 3;
    const uint2 td = tab[type];
    const float decay = __uint_as_float(td.x);
    uint32_t hm = valid ? (uint32_t)(recent >> td.y) & kmask : 0u;
    if (young && hm) hm &= since_mask(cstep[meta_coh(h.meta)], 64u - td.y, t0);
    hit = hm;
    int k1 = -1;
    float c1 = 0.0f;
    if (hm) {
        k1 = __ffs(hm) - 1;
        // changePSR's ((W / decay) * u) * r, or W / decay for Spiking synapses
        c1 = SK == SK_DS ? (c.Wd * c.f.x) * c.f.y : c.Wd;
        hm &= hm - 1;
    }
    float p = valid ? psr : 0.0f;
    if (__ballot(hm != 0) == 0ull) {
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (BG_K_EXACT(KMAX) || k < K) {
                p = (k == k1) ? p + c1 : p;
                p = p * decay;
                acc[k] += p;
            }
        }
    } else {
        // a second delivery within the block: the later ones from a private copy of the state
        ColdDS cd;
        cd.valid = true;
        bool ru_changed = false;
        if (hm) {
            cd.W = ld_off(n.W, h.slot << 2);
            cd.u = c.f.x;
            cd.r = c.f.y;
            cd.last = t0 + (uint64_t)k1;
        }
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (BG_K_EXACT(KMAX) || k < K) {
                p = (k == k1) ? p + c1 : p;
                if ((hm >> k) & 1u)
                    p = p + psr_increment_regs<SK>(n, tr, h.slot, type, t0 + (uint64_t)k, decay, cd, ru_changed);
                p = p * decay;
                acc[k] += p;
            }
        }
    }
    psr = p;
}

// The row of one destination in passes of 64 ranks (lane l takes ranks l, l+64, ...), in a three
// stage software pipeline: while pass p is processed, pass p+1 has its sources' fire words and its
// cold state in flight and pass p+2 its meta and psr. Every load is unconditional, so the waits
// before a pass cover only the loads it needs.
template <int SK, int KMAX, bool DENSE, class Prologue>
__device__ __forceinline__ void impl_row(const DevNet &n, const TypeRegs &tr, const uint2 *tab, const uint64_t *cstep,
                                         const uint64_t *rec, const float2 *fac, uint32_t row, const uint32_t *imap,
                                         uint32_t cnt, uint64_t t0, int K, bool young, float (&acc)[KMAX],
                                         Prologue &&prologue)
{
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const uint32_t P = (cnt + BG_WAVE - 1) / BG_WAVE;
    const uint32_t last = cnt - 1;
    auto rank = [&](uint32_t pass) {
        const uint32_t q = pass * BG_WAVE + (uint32_t)lane;
        return q < cnt ? q : last;
    };
    ImplRank x, y, z;
    ImplCold cx, cy, cz;
    uint64_t fx, fy, fz;
    impl_hot_load<DENSE>(n, row, imap, rank(0), x);
    impl_hot_load<DENSE>(n, row, imap, rank(1), y);
    // keep y's loads older than x's gathers, as in every pass, so that the loop header's wait for
    // the first pass covers the same loads whether it is entered here or from the loop's end
    __builtin_amdgcn_sched_barrier(0);
    fx = ld_off(rec, (x.meta & BG_META_SRC_MASK) << 3);
    impl_cold_load<SK>(n, fac, x, cx);
    // work of the caller's that needs none of the row, done while the first passes' loads fly
    prologue();
    // full passes (every lane has a rank of the row) and, when the row length is not a multiple
    // of 64, a last partial pass in which the lanes past its end add +0.0
    const uint32_t PF = cnt / BG_WAVE;
    auto pass = [&](uint32_t p, ImplRank &cur, ImplCold &ccur, uint64_t &fcur, ImplRank &nxt, ImplCold &cnxt,
                    uint64_t &fnxt, ImplRank &far) {
        impl_hot_load<DENSE>(n, row, imap, rank(p + 2), far);
        fnxt = ld_off(rec, (nxt.meta & BG_META_SRC_MASK) << 3);
        impl_cold_load<SK>(n, fac, nxt, cnxt);
        uint32_t hit = 0;
        float ps = cur.psr;
        if (p < PF) {
            impl_synapse<SK, KMAX, false>(n, tr, tab, cstep, cur, ps, fcur, ccur, true, t0, K, young, acc, hit);
            st_off(n.psr, cur.slot << 2, ps);
        } else {
            const bool valid = p * BG_WAVE + (uint32_t)lane < cnt;
            impl_synapse<SK, KMAX, true>(n, tr, tab, cstep, cur, ps, fcur, ccur, valid, t0, K, young, acc, hit);
            if (valid) st_off(n.psr, cur.slot << 2, ps);
        }
    };
    for (uint32_t p = 0; p < P; p += 3) {
        pass(p, x, cx, fx, y, cy, fy, z);
        if (p + 1 >= P) break;
        pass(p + 1, y, cy, fy, z, cz, fz, x);
        if (p + 2 >= P) break;
        pass(p + 2, z, cz, fz, x, cx, fx, y);
    }
}

// The butterfly of wave_sum -- lanes l and l + 32 first, then l and l + 16, ..., l and l + 1, the
// same pairs and therefore the same bits -- with moves that need no address registers: one
// permute across the halves, a swizzle within 32 lanes, then row shifts inside the first row.
// The total is valid in lane 15 only.
__device__ __forceinline__ float wave_sum_lane15(float v)
{
    v = v + __shfl_xor(v, 32, 64);
    v = v + __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));   // lane ^ 16
    v = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xF, 0xF, false));
    v = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xF, 0xF, false));
    v = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xF, 0xF, false));
    v = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, false));
    return v;
}

// One workgroup of W = BG_BLOCK / 64 wavefronts advances W destinations through one block: those
// at positions d0 .. d0 + W - 1 of the network's longest-row-first order (k_lpt_order). Each
// wavefront walks its destination's row (impl_row) and leaves the K summation-point sums in LDS.
// While its row's first loads are in flight it stages what the per-neuron work needs into LDS --
// its neuron's state, summation point and fire word, the neuron's outgoing (cohort, target class)
// states as a source, and on wavefront 0 the block's noise of all W neurons (lanes (w, k), one pass
// of the generator for the workgroup) -- so that nothing after the row waits on memory; after the
// row it applies this block's deliveries to those source states. The work that is serial in the
// steps -- the stimulus input, the neuron update and the next block's first delivery -- then runs
// once for all W neurons on wavefront 0, lane w taking neuron w of the workgroup. Which workgroup and wavefront a destination lands in
// changes no value: each destination's row, sums and neuron update are its own.
template <int NK, int SK, int KMAX>
__device__ __forceinline__ void fused_impl_group(const DevNet &n, int d0, uint64_t t0, int K, int sel)
{
    constexpr int W = BG_BLOCK / BG_WAVE;
    constexpr int NP = 2 * BG_COH_MAX;
    __shared__ float s_acc[W][KMAX];
    __shared__ float s_noise[W][KMAX];
    __shared__ float s_in[W][KMAX];
    __shared__ uint2 s_tab[W][4];
    __shared__ uint64_t s_coh[W][BG_COH_MAX];
    __shared__ float s_pr[W][NP], s_pu[W][NP];
    __shared__ uint64_t s_plast[W][NP];
    __shared__ NeuronRegs s_nr[W];
    __shared__ float s_sp[W];
    __shared__ uint64_t s_rec[W];
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / BG_WAVE));
    // the workgroup's destinations: positions d0 .. d0 + W - 1 of the longest-row-first order
    auto dest = [&](int g) { return g < n.N ? (int)n.lpt[g] : n.N; };
    const int d = __builtin_amdgcn_readfirstlane(dest(d0 + wv));
    const uint64_t *rec = n.rec[sel];
    const uint32_t cm = n.coh_mask;
    const int npair = SK == SK_DS ? 2 * __popc(cm) : 0;
    TypeRegs tr;
    type_regs(n, tr);

    if (d < n.N) {
        const uint32_t cnt = n.syn_count[d], hi = n.slot_hi[d];
        const bool dense = (cnt == hi);
        const uint32_t row = (uint32_t)d * (uint32_t)n.M;
        const uint32_t *imap = n.in_map + (dense ? 0 : n.in_begin[d]);
        if (lane < 4) s_tab[wv][lane] = make_uint2(__float_as_uint(n.decay[lane]), 64u - (uint32_t)n.tdelay[lane]);
        if (lane < BG_COH_MAX) s_coh[wv][lane] = n.coh_step[lane];
        const float2 *fac = n.cfac + (size_t)sel * BG_COH_MAX * 2 * (size_t)n.N;
        const bool young = t0 < n.young_until;
        // lane j < npair: outgoing state j of this neuron as a source, (cohort j / 2, class j % 2)
        int pc = 0;
        size_t pidx = 0;
        if (SK == SK_DS && lane < npair) {
            uint32_t m = cm;
            for (int j = 0; j < (lane >> 1); j++) m &= m - 1;
            pc = __ffs(m) - 1;
            pidx = ((size_t)pc * 2 + (size_t)(lane & 1)) * (size_t)n.N + (size_t)d;
        }
        auto stage = [&] {
            // loads first, then the noise while they fly, then the LDS copies
            NeuronRegs s;
            float spv = 0.0f, pr = 0.0f, pu = 0.0f;
            uint64_t rv = 0, pl = 0;
            if (lane == 0) {
                neuron_load<NK>(n, d, s);
                spv = n.sp[d];
                rv = rec[d];
            }
            if (SK == SK_DS && lane < npair) {
                pr = n.cst_r[pidx];
                pu = n.cst_u[pidx];
                pl = n.cst_last[pidx];
            }
            // wavefront 0: the block's noise of the workgroup's W neurons, lane j = (w, k), unless
            // neuron w stays refractory through the block (then neuron_advance never reads it)
            if (wv == 0) {
                for (int j = lane; j < W * K; j += BG_WAVE) {
                    const int w = j / K, k = j - w * K;
                    const int dw = dest(d0 + w);
                    if (dw < n.N && n.refr[dw] < K)
                        s_noise[w][k] = bg_noise_normal(n.seed, (uint32_t)dw, t0 + (uint64_t)k);
                }
            }
            if (lane == 0) {
                s_nr[wv] = s;
                s_sp[wv] = spv;
                s_rec[wv] = rv;
            }
            if (SK == SK_DS && lane < npair) {
                s_pr[wv][lane] = pr;
                s_pu[wv][lane] = pu;
                s_plast[wv][lane] = pl;
            }
        };
        float acc[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; k++) acc[k] = 0.0f;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (cnt) {
            if (dense) impl_row<SK, KMAX, true>(n, tr, s_tab[wv], s_coh[wv], rec, fac, row, imap, cnt, t0, K, young, acc, stage);
            else impl_row<SK, KMAX, false>(n, tr, s_tab[wv], s_coh[wv], rec, fac, row, imap, cnt, t0, K, young, acc, stage);
        } else {
            stage();
        }
#pragma unroll
        for (int k = 0; k < KMAX; k++)
            if (BG_K_EXACT(KMAX) || k < K) acc[k] = wave_sum_lane15(acc[k]);
        if (lane == 15) {
#pragma unroll
            for (int k = 0; k < KMAX; k++)
                if (BG_K_EXACT(KMAX) || k < K) s_acc[wv][k] = acc[k];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // this block's deliveries to the neuron's outgoing synapses, from its fire word before the
        // block
        if (SK == SK_DS && lane < npair) {
            float pr = s_pr[wv][lane], pu = s_pu[wv][lane];
            uint64_t plast = s_plast[wv][lane];
            const uint64_t pE = s_coh[wv][pc];
            const int type = (n.ntype[d] == 1 ? 0 : 2) + (lane & 1);
            const uint32_t delay = (uint32_t)n.tdelay[type];
            const float U = pick4(tr.U, type), D = pick4(tr.D, type), F = pick4(tr.F, type);
            const uint32_t kmK = K >= 32 ? 0xFFFFFFFFu : ((1u << K) - 1u);
            uint32_t hr = (uint32_t)(s_rec[wv] >> (64u - delay)) & kmK & since_mask(pE, delay, t0);
            if (hr) {
                do {
                    const int k = __ffs(hr) - 1;
                    hr &= hr - 1;
                    ds_deliver(n, pr, pu, plast, t0 + (uint64_t)k, U, D, F);
                } while (hr);
                n.cst_r[pidx] = pr;
                n.cst_u[pidx] = pu;
                n.cst_last[pidx] = plast;
                s_pr[wv][lane] = pr;
                s_pu[wv][lane] = pu;
                s_plast[wv][lane] = plast;
            }
        }
    }
    __syncthreads();
    if (wv != 0) return;

    uint32_t mask = 0;
    uint64_t rec_new = 0;
    const int dn = lane < W ? dest(d0 + lane) : n.N;
    if (dn < n.N) {
        uint32_t inbits = 0;
        if (n.in_kind != BG_INPUT_NONE) {
            for (int k = 0; k < K; k++) {
                float add;
                if (input_step(n, dn, t0 + (uint64_t)k, &add)) {
                    s_in[lane][k] = add;
                    inbits |= 1u << k;
                }
            }
        }
        NeuronRegs s = s_nr[lane];
        float sp = s_sp[lane];
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (BG_K_EXACT(KMAX) || k < K) {
                if ((inbits >> k) & 1u) sp = sp + s_in[lane][k];
                if (neuron_advance<NK>(n, dn, t0 + (uint64_t)k, sp, s_noise[lane][k], s)) mask |= 1u << k;
                sp = 0.0f + s_acc[lane][k];
            }
        }
        neuron_store<NK>(n, dn, s);
        n.sp[dn] = sp;
        n.mask[sel ^ 1][dn] = mask;
        if (mask) n.fired_step[dn] = t0 + (uint64_t)(31 - __clz(mask));
        rec_new = (s_rec[lane] >> K) | ((uint64_t)mask << (64 - K));
        n.rec[sel ^ 1][dn] = rec_new;
    }
    // the first delivery of the next block (of up to `look` steps) to the outgoing synapses of
    // neuron w, lane w * npair + j
    if (SK == SK_DS && npair) {
        const int w = lane / npair, j = lane - w * npair;
        const uint64_t rw = __shfl(rec_new, w < W ? w : 0);
        const int dd = __shfl(dn, w < W ? w : 0);
        if (w < W && dd < n.N) {
            uint32_t m = cm;
            for (int q = 0; q < (j >> 1); q++) m &= m - 1;
            const int pc = __ffs(m) - 1;
            const int tsub = j & 1;
            const int type = (n.ntype[dd] == 1 ? 0 : 2) + tsub;
            const uint32_t delay = (uint32_t)n.tdelay[type];
            const uint64_t pE = s_coh[0][pc];
            const uint64_t T = t0 + (uint64_t)K;
            const uint32_t kmL = n.look >= 32 ? 0xFFFFFFFFu : ((1u << n.look) - 1u);
            const uint32_t hl = (uint32_t)(rw >> (64u - delay)) & kmL & since_mask(pE, delay, T);
            if (hl) {
                const float U = pick4(tr.U, type), D = pick4(tr.D, type), F = pick4(tr.F, type);
                float r2 = s_pr[w][j], u2 = s_pu[w][j];
                uint64_t l2 = s_plast[w][j];
                ds_deliver(n, r2, u2, l2, T + (uint64_t)(__ffs(hl) - 1), U, D, F);
                n.cfac[(((size_t)(sel ^ 1) * BG_COH_MAX + (size_t)pc) * 2 + (size_t)tsub) * (size_t)n.N + (size_t)dd] =
                    make_float2(u2, r2);
            }
        }
    }
}

template <int NK, int SK, int KMAX>
__global__ __launch_bounds__(BG_BLOCK) __attribute__((amdgpu_waves_per_eu(BG_IMPL_WAVES))) void k_fused_impl(
    DevNet n, uint64_t t0, int K, int sel)
{
    fused_impl_group<NK, SK, KMAX>(n, (int)(blockIdx.x * (BG_BLOCK / BG_WAVE)), t0, K, sel);
}

// Many networks: workgroup q advances the four destinations at positions quads[q].y .. + 3 of
// member quads[q].x (members are padded to whole workgroups, so a workgroup's network is uniform).
template <int NK, int SK, int KMAX>
__global__ __launch_bounds__(BG_BLOCK) __attribute__((amdgpu_waves_per_eu(BG_IMPL_WAVES))) void k_fused_impl_batch(
    BatchMap bm, int K, int Kfull, int launch)
{
    const uint32_t q = blockIdx.x;
    if (q >= bm.nquads) return;
    const uint2 e = bm.quads[q];
    const BatchNet &b = bm.nets[e.x];
    const uint64_t t0 = b.t0 + (uint64_t)launch * (uint64_t)Kfull;
    const int sel = b.sel ^ (launch & 1);
    fused_impl_group<NK, SK, KMAX>(b.n, (int)e.y, t0, K, sel);
}

// The explicit slot state an implicit network stands for at step boundary T (fire words in
// rec[sel]): each synapse's 32-bit ring in absolute orientation (bit (s + delay) mod 32 for each
// spike s of its source in [T - delay, T) at or after its creation) and, for DS, its cohort's r, u,
// lastSpike. One thread per slot.
static __global__ __launch_bounds__(BG_BLOCK) void k_materialize(DevNet n, uint64_t T, int sel)
{
    const uint64_t slot = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (slot >= n.S) return;
    const uint32_t meta = n.meta[slot];
    if (!(meta & BG_META_IN_USE)) return;
    const uint32_t src = meta & BG_META_SRC_MASK;
    const int type = (meta >> BG_META_TYPE_SHIFT) & 3;
    const uint32_t delay = (meta >> BG_META_DELAY_SHIFT) & 0xFF;
    const int c = meta_coh(meta);
    const uint64_t E = n.coh_step[c];
    const uint64_t rec = n.rec[sel][src];
    uint32_t ring = 0;
    for (uint32_t j = 1; j <= delay && j <= 64; j++) {
        if (T < (uint64_t)j) break;
        const uint64_t s = T - j;
        if (s < E) break;
        if ((rec >> (64u - j)) & 1ull) ring |= 1u << ((uint32_t)(s + delay) & 31u);
    }
    n.ring[slot] = ring;
    if (n.cst_r) {
        const size_t i = ((size_t)c * 2 + (size_t)(type & 1)) * (size_t)n.N + src;
        n.r[slot] = n.cst_r[i];
        n.us[slot] = n.cst_u[i];
        n.last[slot] = n.cst_last[i];
    }
}

// A new cohort c created at `step`: its creation step, and for DS the createSynapse state
// (r0, u0, no spike yet) for every (target class, source). Threads cover 2 N entries.
static __global__ void k_coh_init(DevNet n, int c, uint64_t step, float r0, float u0)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i == 0) n.coh_step[c] = step;
    if (!n.cst_r || i >= 2 * (uint64_t)n.N) return;
    const size_t k = (size_t)c * 2 * (size_t)n.N + i;
    n.cst_r[k] = r0;
    n.cst_u[k] = u0;
    n.cst_last[k] = BG_NO_SPIKE;
}

// Cohort ids of the synapses in use (bit c of *out), over a grid-strided pass of the slots.
static __global__ __launch_bounds__(BG_BLOCK) void k_coh_usage(DevNet n, uint32_t *out)
{
    __shared__ uint32_t s_bits;
    if (threadIdx.x == 0) s_bits = 0;
    __syncthreads();
    uint32_t bits = 0;
    for (uint64_t slot = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; slot < n.S;
         slot += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t meta = n.meta[slot];
        if (meta & BG_META_IN_USE) bits |= 1u << meta_coh(meta);
    }
    for (int off = 32; off >= 1; off >>= 1) bits |= (uint32_t)__shfl_xor((int)bits, off);
    if ((threadIdx.x & (BG_WAVE - 1)) == 0 && bits) atomicOr(&s_bits, bits);
    __syncthreads();
    if (threadIdx.x == 0 && s_bits) atomicOr(out, s_bits);
}
