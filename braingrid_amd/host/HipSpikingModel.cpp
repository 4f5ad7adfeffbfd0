// HipSpikingModel: the host half of the MI355X model (include/bghost.h).
//
// Mirrors the reference's IModel life cycle (Core/IModel.h:42-152) over the C ABI of
// include/bghip.h:
//   setupSim          Model::setupSim (Core/Model.cpp:136-158): layout, createAllNeurons with the
//                     global MTRand rng(1) (Utils/Global.cpp:79), ConnGrowth radii; then the GPU
//                     arena is created and uploaded (GPUSpikingModel::setupSim,
//                     Core/GPUSpikingModel.cu:106-145);
//   advance           GPUSpikingModel::advance (Core/GPUSpikingModel.cu:188-225);
//   updateConnections GPUSpikingModel::updateConnections (:246-259), entirely on the GPU;
//   updateHistory     XmlGrowthRecorder::compileHistories (Recorders/XmlGrowthRecorder.cpp:72-97)
//                     from downloaded spikes, rates and radii, then the device clamp/clear;
//   saveData          XmlGrowthRecorder::saveSimData (:104-154), byte-compatible output.
// Host arithmetic keeps the reference's float/double types, compiled with -ffp-contract=off.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <new>
#include <string>
#include <thread>
#include <vector>
#include "../../include/bghost.h"

namespace bg {

// MT19937 exactly as MTRand (RNG/MersenneTwister.cpp:66-291), restated from the MT19937 paper.
class MTRand {
  public:
    explicit MTRand(uint32_t seed) { reseed(seed); }
    void reseed(uint32_t seed)
    {
        s_[0] = seed;
        for (int i = 1; i < 624; i++) s_[i] = 1812433253u * (s_[i - 1] ^ (s_[i - 1] >> 30)) + (uint32_t)i;
        reload();
    }
    uint32_t randInt()
    {
        if (left_ == 0) reload();
        --left_;
        uint32_t y = s_[next_++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        return y ^ (y >> 18);
    }
    float rand() { return (float)((double)(float)randInt() * (1.0 / 4294967295.0)); }
    float inRange(float lo, float hi)
    {
        float v = rand();
        float range = hi - lo;
        v *= range;
        v += lo;
        return v;
    }

  private:
    static uint32_t twist(uint32_t m, uint32_t a, uint32_t b)
    {
        return m ^ (((a & 0x80000000u) | (b & 0x7fffffffu)) >> 1) ^ ((uint32_t)(-(int32_t)(b & 1u)) & 0x9908b0dfu);
    }
    void reload()
    {
        int i = 0;
        for (; i < 227; i++) s_[i] = twist(s_[i + 397], s_[i], s_[i + 1]);
        for (; i < 623; i++) s_[i] = twist(s_[i - 227], s_[i], s_[i + 1]);
        s_[623] = twist(s_[396], s_[623], s_[0]);
        left_ = 624;
        next_ = 0;
    }
    uint32_t s_[624];
    int left_ = 0, next_ = 0;
};

enum { INH = 1, EXC = 2 };

class HipSpikingModel {
  public:
    HipSpikingModel(bghip_ctx *ctx, const bghost_params &p) : ctx_(ctx), p_(p), rng_(1u) {}
    ~HipSpikingModel() { cleanupSim(); }

    int setupSim();
    int advance(uint64_t n) { return bghip_advance(net_, n); }
    int advanceUntilGrowth()
    {
        uint64_t steps = (uint64_t)(p_.epoch_duration / p_.dt);   // Core/Simulator.cpp:159-160
        return advance(steps);
    }
    int updateConnections(bghip_growth_stats *st)
    {
        if (p_.conn_kind != BGHOST_CONN_GROWTH) return BGHIP_OK;
        return bghip_update_connections(net_, st);
    }
    int updateHistory();
    int syncRecorder();
    int serialize(const char *path);
    int deserialize(const char *path);
    int saveData(const char *path);
    void cleanupSim()
    {
        if (net_) bghip_net_destroy(net_);
        net_ = nullptr;
    }
    int setRadii(float r)
    {
        int rc = syncRecorder();
        if (rc) return rc;
        for (auto &x : radii_) x = r;
        return bghip_upload_growth(net_, &growth_, radii_.data(), rates_.data());
    }
    const void *array(const char *name, int64_t *count);
    bghip_net *net() { return net_; }
    int epoch() const { return epoch_; }

  private:
    void setupLayout();
    void createAllNeurons();
    int setupStaticConnections();
    int constructInput();
    int initInput();
    bghip_synapse_types synapseTypes() const;

    bghip_ctx *ctx_;
    bghost_params p_;
    MTRand rng_;
    bghip_net *net_ = nullptr;
    int N_ = 0, max_spikes_ = 0, epoch_ = 0;
    bool recorder_stale_ = false;   // device histories newer than the host copies
    std::vector<float> xloc_, yloc_;
    std::vector<int32_t> ntype_;
    std::vector<uint8_t> starter_;
    std::vector<float> Vm_, Vthresh_, Vreset_, Vrest_, Vinit_, I0_, Inoise_, Iinject_, C1_, C2_, Trefract_, Tau_, Rm_;
    std::vector<float> A_, B_, C_, D_, u_, C3_;
    std::vector<float> radii_, rates_, radiiHistory_, ratesHistory_, burst_, spikesHist_;
    std::vector<float> in_values_;     // regular input: per-neuron pulse value
    std::vector<int32_t> in_shift_;    // regular input: per-neuron cycle shift
    std::vector<int32_t> spikeCount_, spikeOffset_;
    std::vector<uint64_t> hist_;
    bghip_growth growth_;
};

void HipSpikingModel::setupLayout()
{
    const int N = N_;
    for (int i = 0; i < N; i++) {   // Layouts/Layout.cpp:155-157, 104-106, 117-119
        xloc_[i] = (float)(i % p_.width);
        yloc_[i] = (float)(i / p_.width);
        ntype_[i] = EXC;
        starter_[i] = 0;
    }
    if (p_.layout_kind == BGHOST_LAYOUT_FIXED) {   // Layouts/FixedLayout.cpp:247-279
        for (int i = 0; i < p_.n_inh; i++) ntype_[p_.inh_list[i]] = INH;
        for (int i = 0; i < p_.n_active; i++) starter_[p_.active_list[i]] = 1;
    } else {                                        // Layouts/DynamicLayout.cpp:98-165
        int num_exc = (int)((double)(p_.frac_exc * N) + 0.5);
        int num_inh = N - num_exc;
        std::vector<int> rg(num_inh > 0 ? num_inh : 1);
        for (int i = 0; i < num_inh; i++) rg[i] = i;
        for (int i = num_inh; i < N; i++) {
            int j = (int)(rng_.rand() * N);
            if (j < num_inh) rg[j] = i;
        }
        for (int i = 0; i < num_inh; i++) ntype_[rg[i]] = INH;
        uint32_t want = (uint32_t)((double)(p_.frac_starter * N) + 0.5), got = 0;
        while (got < want) {
            int i = (int)rng_.inRange(0.0f, (float)N);
            if (ntype_[i] == EXC && !starter_[i]) {
                starter_[i] = 1;
                got++;
            }
        }
    }
}

int HipSpikingModel::constructInput()
{   // SInputRegular's constructor (Inputs/SInputRegular.cpp:26-110): the <Values> table tiled over
    // the grid and, for sync "no", one shift per neuron drawn from the global rng; "wave" shifts by
    // column. SInputPoisson draws nothing here.
    if (p_.input_kind != BGHIP_INPUT_REGULAR) return BGHIP_OK;
    const int W = p_.width, H = p_.height;
    const float dt = p_.dt;
    const int dur = (int)((double)(p_.input_duration / dt) + 0.5);
    const int itv = (int)((double)(p_.input_interval / dt) + 0.5);
    const int cycle = dur + itv;
    if (cycle <= 0) return BGHIP_EINVAL;
    in_values_.assign((size_t)N_, 0.0f);
    in_shift_.assign((size_t)N_, 0);
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) in_values_[(size_t)i * W + j] = p_.input_values[(i % 10) * 10 + j % 10];
    if (p_.input_sync == BGHOST_SYNC_NO) {
        for (int i = 0; i < H; i++)
            for (int j = 0; j < W; j++) in_shift_[(size_t)i * W + j] = (int)rng_.inRange(0.0f, (float)(cycle - 1));
    } else if (p_.input_sync == BGHOST_SYNC_WAVE) {
        for (int i = 0; i < H; i++)
            for (int j = 0; j < W; j++) in_shift_[(size_t)i * W + j] = (cycle / W) * j;
    } else if (p_.input_sync != BGHOST_SYNC_YES) {
        return BGHIP_EINVAL;
    }
    return BGHIP_OK;
}

int HipSpikingModel::initInput()
{   // ISInput::init on the device (Inputs/SInputPoisson.cpp:124-146, Inputs/SInputRegular.cpp)
    if (p_.input_kind == BGHIP_INPUT_NONE) return BGHIP_OK;
    bghip_input in;
    memset(&in, 0, sizeof in);
    in.kind = p_.input_kind;
    std::vector<uint8_t> mask;
    if (p_.input_kind == BGHIP_INPUT_POISSON) {
        in.fr_mean = p_.input_fr_mean;
        in.weight = p_.input_weight;
        in.neuron_type = ntype_.data();
        if (p_.n_input_mask >= 0) {
            mask.assign((size_t)N_, 0);
            for (int k = 0; k < p_.n_input_mask; k++) {
                const int i = p_.input_mask[k];
                if (i < 0 || i >= N_) return BGHIP_EINVAL;
                mask[i] = 1;
            }
            in.mask = mask.data();
        }
    } else if (p_.input_kind == BGHIP_INPUT_REGULAR) {
        in.duration = p_.input_duration;
        in.interval = p_.input_interval;
        in.values = in_values_.data();
        in.shift = in_shift_.data();
    } else {
        return BGHIP_EINVAL;
    }
    return bghip_upload_input(net_, &in, nullptr);
}

namespace {
struct DistDest {   // the record ConnStatic sorts: ordered by distance alone (Connections/ConnStatic.h:138-146)
    float dist;
    int dest;
    bool operator<(const DistDest &o) const { return dist < o.dist; }
};
}  // namespace

int HipSpikingModel::setupStaticConnections()
{   // ConnStatic::setupConnections (Connections/ConnStatic.cpp:33-88). For each source in ascending
    // order: the other neurons within threshConnsRadius, std::sort by distance (so equal distances
    // end in the standard library's order, as in the reference), the first nConnsPerNeuron of them
    // connected, each weight drawn from the global rng right after its synapse is added. The
    // candidate lists of different sources are independent, so they are built and sorted on worker
    // threads; the weight draws then run in the reference's order. Distances are Layout::dist: the
    // float square root of the float squared distance (Layouts/Layout.cpp:54-69).
    const int N = N_, k = p_.static_nconns;
    const float thresh = p_.static_thresh;
    if (k < 0 || !(thresh >= 0)) return BGHIP_EINVAL;
    std::vector<std::vector<DistDest>> chosen(N);
    auto work = [&](int t, int nt) {
        std::vector<DistDest> cand;
        cand.reserve(N);
        for (int s = t; s < N; s += nt) {
            cand.clear();
            for (int d = 0; d < N; d++) {
                if (d == s) continue;
                const float dx = xloc_[s] - xloc_[d], dy = yloc_[s] - yloc_[d];
                const float r = sqrtf(dx * dx + dy * dy);
                if (r <= thresh) cand.push_back(DistDest{r, d});
            }
            std::sort(cand.begin(), cand.end());
            chosen[s].assign(cand.begin(), cand.begin() + std::min<size_t>(cand.size(), (size_t)k));
        }
    };
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work, t, nt);
    work(0, nt);
    for (auto &th : pool) th.join();
    std::vector<int32_t> src, dst, indeg(N, 0);
    std::vector<float> w;
    for (int s = 0; s < N; s++)
        for (const DistDest &c : chosen[s]) {
            const int d = c.dest;
            const bool exc_src = ntype_[s] == EXC;   // synSign(synType(s, d)) > 0: EE and EI
            const float *r = exc_src ? p_.static_exc_w : p_.static_inh_w;
            src.push_back(s);
            dst.push_back(d);
            w.push_back(rng_.inRange(r[0], r[1]));
            // createSynapseImap asserts synapse_counts < maxSynapsesPerNeuron (Synapses/AllSynapses.cpp:231)
            if (++indeg[d] >= p_.max_synapses_per_neuron) return BGHIP_EINVAL;
        }
    return bghip_add_synapses(net_, (int64_t)src.size(), src.data(), dst.data(), w.data());
}

void HipSpikingModel::createAllNeurons()
{   // AllIFNeurons::createNeuron (Neurons/AllIFNeurons.cpp:398-503), AllIZHNeurons (:413-476)
    const float dt = p_.dt;
    const bool izh = p_.neuron_model == BGHIP_NEURON_IZH;
    for (int i = 0; i < N_; i++) {
        Rm_[i] = (float)1e6;
        Tau_[i] = (float)(3e-8 * 1e6);
        Iinject_[i] = rng_.inRange(p_.Iinject[0], p_.Iinject[1]);
        Inoise_[i] = rng_.inRange(p_.Inoise[0], p_.Inoise[1]);
        Vthresh_[i] = rng_.inRange(p_.Vthresh[0], p_.Vthresh[1]);
        Vrest_[i] = rng_.inRange(p_.Vresting[0], p_.Vresting[1]);
        Vreset_[i] = rng_.inRange(p_.Vreset[0], p_.Vreset[1]);
        Vinit_[i] = rng_.inRange(p_.Vinit[0], p_.Vinit[1]);
        Vm_[i] = Vinit_[i];
        if (Tau_[i] > 0) {
            C1_[i] = expf(-dt / Tau_[i]);
            C2_[i] = Rm_[i] * (1 - C1_[i]);
        } else {
            C1_[i] = 0.0f;
            C2_[i] = Rm_[i];
        }
        I0_[i] = Iinject_[i] + Vrest_[i] / Rm_[i];
        if (izh) C3_[i] = dt * 1000;
        Trefract_[i] = ntype_[i] == INH ? (float)2.0e-3 : (float)3.0e-3;
        if (starter_[i]) {
            Vthresh_[i] = rng_.inRange(p_.starter_Vthresh[0], p_.starter_Vthresh[1]);
            Vreset_[i] = rng_.inRange(p_.starter_Vreset[0], p_.starter_Vreset[1]);
            Trefract_[i] = (float)3.0e-3;
        }
        if (izh) {
            const float *a = ntype_[i] == EXC ? p_.excA : p_.inhA, *b = ntype_[i] == EXC ? p_.excB : p_.inhB;
            const float *c = ntype_[i] == EXC ? p_.excC : p_.inhC, *d = ntype_[i] == EXC ? p_.excD : p_.inhD;
            A_[i] = rng_.inRange(a[0], a[1]);
            B_[i] = rng_.inRange(b[0], b[1]);
            C_[i] = rng_.inRange(c[0], c[1]);
            D_[i] = rng_.inRange(d[0], d[1]);
            u_[i] = 0.0f;
        }
    }
}

bghip_synapse_types HipSpikingModel::synapseTypes() const
{   // per-type constants of createSynapse (Synapses/AllSpikingSynapses.cpp:191-233,
    // AllDSSynapses.cpp:180-218, AllSTDPSynapses.cpp:252-273)
    bghip_synapse_types t;
    memset(&t, 0, sizeof t);
    const float tau[4] = {(float)6e-3, (float)6e-3, (float)3e-3, (float)3e-3};
    const float delay[4] = {(float)0.8e-3, (float)0.8e-3, (float)0.8e-3, (float)1.5e-3};
    const float U[4] = {(float)0.32, (float)0.25, (float)0.05, (float)0.5};
    const float D[4] = {(float)0.144, (float)0.7, (float)0.125, (float)1.1};
    const float F[4] = {(float)0.06, (float)0.02, (float)1.2, (float)0.05};
    for (int k = 0; k < 4; k++) {
        t.tau[k] = tau[k];
        t.decay[k] = expf(-p_.dt / tau[k]);
        t.total_delay[k] = (int)(delay[k] / p_.dt) + 1;
        t.U[k] = U[k];
        t.D[k] = D[k];
        t.F[k] = F[k];
    }
    t.u0 = (float)0.4;
    t.r0 = 1.0f;
    t.Apos = 0.5f; t.Aneg = -0.5f; t.STDPgap = (float)2e-3; t.tauspost = (float)75e-3; t.tauspre = (float)34e-3;
    t.taupos = (float)15e-3; t.tauneg = (float)35e-3; t.Wex = 1.0f; t.mupos = 0.0f; t.muneg = 0.0f;
    t.useFroemkeDanSTDP = 1;
    return t;
}

int HipSpikingModel::setupSim()
{
    N_ = p_.width * p_.height;
    if (N_ <= 0) return BGHIP_EINVAL;
    max_spikes_ = (int)(p_.epoch_duration * p_.max_firing_rate);
    const int E = p_.num_epochs > 0 ? p_.num_epochs : 1;
    size_t N = N_;
    for (auto *v : {&xloc_, &yloc_, &Vm_, &Vthresh_, &Vreset_, &Vrest_, &Vinit_, &I0_, &Inoise_, &Iinject_, &C1_,
                    &C2_, &Trefract_, &Tau_, &Rm_, &A_, &B_, &C_, &D_, &u_, &C3_, &radii_, &rates_})
        v->assign(N, 0.0f);
    ntype_.assign(N, EXC);
    starter_.assign(N, 0);
    spikeCount_.assign(N, 0);
    spikeOffset_.assign(N, 0);
    hist_.assign(N * (size_t)max_spikes_, ~0ull);
    radiiHistory_.assign((size_t)(E + 1) * N, 0.0f);
    ratesHistory_.assign((size_t)(E + 1) * N, 0.0f);
    burst_.assign((size_t)(int)(p_.epoch_duration * p_.num_epochs), 0.0f);
    spikesHist_.assign((size_t)(int)(p_.epoch_duration * p_.num_epochs * 100), 0.0f);
    rng_.reseed(1u);
    int rc0 = constructInput();   // the stimulus object exists before Simulator::setup
    if (rc0) return rc0;
    setupLayout();
    createAllNeurons();
    for (size_t i = 0; i < N; i++) {   // ConnGrowth::setupConnections, recorder initDefaultValues
        radii_[i] = p_.start_radius;
        radiiHistory_[i] = p_.start_radius;
    }
    growth_.epsilon = p_.epsilon; growth_.beta = p_.beta; growth_.rho = p_.rho; growth_.targetRate = p_.target_rate;
    growth_.maxRate = p_.target_rate / p_.epsilon; growth_.minRadius = p_.min_radius;
    growth_.startRadius = p_.start_radius; growth_.epochDuration = p_.epoch_duration;

    bghip_net_desc d;
    memset(&d, 0, sizeof d);
    d.num_neurons = N_;
    d.max_synapses_per_neuron = p_.max_synapses_per_neuron;
    d.neuron_model = p_.neuron_model;
    d.synapse_model = p_.synapse_model;
    d.ring_bits = p_.ring_bits ? p_.ring_bits : 32;
    d.max_spikes = max_spikes_;
    d.dt = p_.dt;
    d.noise_seed = p_.seed;
    d.block_steps = p_.block_steps;
    d.types = synapseTypes();
    int rc = bghip_net_create(ctx_, &d, &net_);
    if (rc) return rc;
    std::vector<float> zero(N, 0.0f);
    std::vector<int32_t> izero(N, 0);
    bghip_neurons h;
    memset(&h, 0, sizeof h);
    h.Vm = Vm_.data(); h.Vthresh = Vthresh_.data(); h.Vreset = Vreset_.data(); h.I0 = I0_.data();
    h.Inoise = Inoise_.data(); h.C1 = C1_.data(); h.C2 = C2_.data(); h.Trefract = Trefract_.data();
    h.summation_map = zero.data(); h.nStepsInRefr = izero.data(); h.spikeCount = spikeCount_.data();
    h.spikeCountOffset = spikeOffset_.data(); h.spike_history = hist_.data();
    if (p_.neuron_model == BGHIP_NEURON_IZH) {
        h.Aconst = A_.data(); h.Bconst = B_.data(); h.Cconst = C_.data(); h.Dconst = D_.data(); h.u = u_.data();
        h.C3 = C3_.data();
    }
    if ((rc = bghip_upload_neurons(net_, &h))) return rc;
    if ((rc = bghip_upload_layout(net_, xloc_.data(), yloc_.data(), ntype_.data()))) return rc;
    if ((rc = bghip_upload_growth(net_, &growth_, radii_.data(), rates_.data()))) return rc;
    // m_conns->setupConnections (Core/Model.cpp:154): ConnGrowth starts without synapses
    if (p_.conn_kind == BGHOST_CONN_STATIC && (rc = setupStaticConnections())) return rc;
    if ((rc = initInput())) return rc;   // pInput->init after the model's setup
    // the recorder's histories live on the device (XmlRecorder::initDefaultValues and
    // XmlGrowthRecorder::initDefaultValues give row 0 and zero bins)
    const bool growth = p_.conn_kind == BGHOST_CONN_GROWTH;
    return bghip_recorder_enable(net_, (int64_t)burst_.size(), (int64_t)spikesHist_.size(), E, growth ? 1 : 0,
                                 ratesHistory_.data(), radiiHistory_.data());
}

int HipSpikingModel::updateHistory()
{   // XmlGrowthRecorder::compileHistories (Recorders/XmlGrowthRecorder.cpp:72-97) and
    // clearSpikeCounts, both on the device; the host copies are refreshed when read
    int rc = bghip_update_history(net_);
    if (rc) return rc;
    epoch_++;
    recorder_stale_ = true;
    return BGHIP_OK;
}

int HipSpikingModel::syncRecorder()
{
    if (!recorder_stale_) return BGHIP_OK;
    const bool growth = p_.conn_kind == BGHOST_CONN_GROWTH;
    int rc = bghip_recorder_download(net_, burst_.data(), spikesHist_.data(), growth ? ratesHistory_.data() : nullptr,
                                     growth ? radiiHistory_.data() : nullptr);
    if (!rc && growth) rc = bghip_download_growth(net_, radii_.data(), rates_.data());
    if (!rc) recorder_stale_ = false;
    return rc;
}

int HipSpikingModel::serialize(const char *path)
{   // Model::serialize (Core/Model.cpp:85-98): the whole device network, recorder included
    return bghip_net_save(net_, path);
}

int HipSpikingModel::deserialize(const char *path)
{   // Model::deserialize (Core/Model.cpp:62-76), called after setupSim (Core/BGDriver.cpp:88-95):
    // the host arrays setupSim drew stay, the device network is replaced by the checkpoint's,
    // which must have been saved by a model of the same configuration.
    bghip_net *nn = nullptr;
    int rc = bghip_net_load(ctx_, path, &nn);
    if (rc) return rc;
    bghip_net_desc a, b;
    bghip_net_get_desc(net_, &a);
    bghip_net_get_desc(nn, &b);
    int32_t e = -1;
    bghip_recorder_info(nn, nullptr, nullptr, nullptr, nullptr, &e);
    if (a.num_neurons != b.num_neurons || a.max_synapses_per_neuron != b.max_synapses_per_neuron ||
        a.neuron_model != b.neuron_model || a.synapse_model != b.synapse_model || a.max_spikes != b.max_spikes ||
        a.dt != b.dt || a.noise_seed != b.noise_seed || a.ring_bits != b.ring_bits || e < 0) {
        bghip_net_destroy(nn);
        return BGHIP_EINVAL;
    }
    bghip_net_destroy(net_);
    net_ = nn;
    epoch_ = e;
    recorder_stale_ = true;
    return syncRecorder();
}

static void put_matrix(FILE *f, const char *name, const float *v, int rows, int cols)
{   // CompleteMatrix::toXML / Print (Matrix/CompleteMatrix.cpp:225-249): "%g" is ostream's default
    fprintf(f, "   <Matrix name=\"%s\" type=\"complete\" rows=\"%d\" columns=\"%d\" multiplier=\"1.0\">\n   ", name,
            rows, cols);
    for (int r = 0; r < rows; r++) {
        for (int c = 0; c < cols; c++) fprintf(f, "%g ", (double)v[(size_t)r * cols + c]);
        fprintf(f, "\n");
    }
    fprintf(f, "\n</Matrix>\n");
}
static void put_vector(FILE *f, const char *name, const float *v, int n)
{   // VectorMatrix::toXML / Print (Matrix/VectorMatrix.cpp:173-190)
    fprintf(f, "   <Matrix name=\"%s\" type=\"complete\" rows=\"1\" columns=\"%d\" multiplier=\"1.0\">\n   ", name, n);
    for (int i = 0; i < n; i++) fprintf(f, "%g ", (double)v[i]);
    fprintf(f, "\n</Matrix>\n");
}

int HipSpikingModel::saveData(const char *path)
{   // XmlGrowthRecorder::saveSimData (Recorders/XmlGrowthRecorder.cpp:104-154)
    int rc = syncRecorder();
    if (rc) return rc;
    FILE *f = fopen(path, "w");
    if (!f) return BGHIP_EINVAL;
    const int E = p_.num_epochs > 0 ? p_.num_epochs : 1;
    std::vector<float> types(N_), starters;
    for (int i = 0; i < N_; i++) {
        types[i] = (float)ntype_[i];
        if (starter_[i]) starters.push_back((float)i);
    }
    fprintf(f, "<?xml version=\"1.0\" standalone=\"no\"?>\n<!-- State output file for the DCT growth modeling-->\n");
    fprintf(f, "<SimState>\n");
    if (p_.conn_kind == BGHOST_CONN_GROWTH) {
        put_matrix(f, "radiiHistory", radiiHistory_.data(), E + 1, N_);
        put_matrix(f, "ratesHistory", ratesHistory_.data(), E + 1, N_);
    }
    put_vector(f, "burstinessHist", burst_.data(), (int)burst_.size());
    put_vector(f, "spikesHistory", spikesHist_.data(), (int)spikesHist_.size());
    put_vector(f, "xloc", xloc_.data(), N_);
    put_vector(f, "yloc", yloc_.data(), N_);
    put_vector(f, "neuronTypes", types.data(), N_);
    if (!starters.empty()) put_vector(f, "starterNeurons", starters.data(), (int)starters.size());
    put_vector(f, "neuronThresh", Vthresh_.data(), N_);
    fprintf(f, "   <Matrix name=\"Tsim\" type=\"complete\" rows=\"1\" columns=\"1\" multiplier=\"1.0\">\n");
    fprintf(f, "   %g\n</Matrix>\n", (double)p_.epoch_duration);
    fprintf(f, "   <Matrix name=\"simulationEndTime\" type=\"complete\" rows=\"1\" columns=\"1\" multiplier=\"1.0\">\n");
    fprintf(f, "   %g\n</Matrix>\n", (double)((float)bghip_net_step(net_) * p_.dt));
    fprintf(f, "</SimState>\n");
    fclose(f);
    return BGHIP_OK;
}

const void *HipSpikingModel::array(const char *name, int64_t *count)
{
    if (syncRecorder()) return nullptr;
    struct Item { const char *n; const void *p; size_t c; };
    const Item items[] = {
        {"xloc", xloc_.data(), xloc_.size()}, {"yloc", yloc_.data(), yloc_.size()},
        {"neuron_type_map", ntype_.data(), ntype_.size()}, {"starter_map", starter_.data(), starter_.size()},
        {"Vm", Vm_.data(), Vm_.size()}, {"Vthresh", Vthresh_.data(), Vthresh_.size()},
        {"Vreset", Vreset_.data(), Vreset_.size()}, {"Vinit", Vinit_.data(), Vinit_.size()},
        {"Iinject", Iinject_.data(), Iinject_.size()}, {"Inoise", Inoise_.data(), Inoise_.size()},
        {"I0", I0_.data(), I0_.size()}, {"C1", C1_.data(), C1_.size()}, {"C2", C2_.data(), C2_.size()},
        {"Trefract", Trefract_.data(), Trefract_.size()}, {"Aconst", A_.data(), A_.size()},
        {"Bconst", B_.data(), B_.size()}, {"Cconst", C_.data(), C_.size()}, {"Dconst", D_.data(), D_.size()},
        {"C3", C3_.data(), C3_.size()}, {"radii", radii_.data(), radii_.size()},
        {"rates", rates_.data(), rates_.size()}, {"radiiHistory", radiiHistory_.data(), radiiHistory_.size()},
        {"ratesHistory", ratesHistory_.data(), ratesHistory_.size()},
        {"burstinessHist", burst_.data(), burst_.size()}, {"spikesHistory", spikesHist_.data(), spikesHist_.size()},
    };
    for (const Item &it : items)
        if (!strcmp(it.n, name)) {
            if (count) *count = (int64_t)it.c;
            return it.p;
        }
    return nullptr;
}

}  // namespace bg

struct bghost_model {
    bg::HipSpikingModel *m;
    std::vector<int32_t> inh, act, inmask;
};

extern "C" int bghost_model_create(bghip_ctx *ctx, const bghost_params *p, bghost_model **out)
{
    if (!ctx || !p || !out) return BGHIP_EINVAL;
    bghost_model *w = new (std::nothrow) bghost_model();
    if (!w) return BGHIP_ENOMEM;
    bghost_params q = *p;
    w->inh.assign(p->inh_list, p->inh_list + (p->inh_list ? p->n_inh : 0));
    w->act.assign(p->active_list, p->active_list + (p->active_list ? p->n_active : 0));
    w->inmask.assign(p->input_mask, p->input_mask + (p->input_mask && p->n_input_mask > 0 ? p->n_input_mask : 0));
    q.inh_list = w->inh.data();
    q.active_list = w->act.data();
    q.input_mask = w->inmask.data();
    if (!p->input_mask && q.n_input_mask > 0) q.n_input_mask = 0;
    const int N = p->width * p->height;
    for (int v : w->inh)
        if (v < 0 || v >= N) { delete w; return BGHIP_EINVAL; }
    for (int v : w->act)
        if (v < 0 || v >= N) { delete w; return BGHIP_EINVAL; }
    w->m = new (std::nothrow) bg::HipSpikingModel(ctx, q);
    if (!w->m) { delete w; return BGHIP_ENOMEM; }
    int rc = w->m->setupSim();
    if (rc) {
        delete w->m;
        delete w;
        return rc;
    }
    *out = w;
    return BGHIP_OK;
}

extern "C" void bghost_model_destroy(bghost_model *w)
{
    if (!w) return;
    delete w->m;
    delete w;
}
extern "C" bghip_net *bghost_model_net(bghost_model *w) { return w ? w->m->net() : nullptr; }
extern "C" int bghost_model_advance(bghost_model *w, uint64_t n) { return w ? w->m->advance(n) : BGHIP_EINVAL; }
extern "C" int bghost_model_advance_epoch(bghost_model *w) { return w ? w->m->advanceUntilGrowth() : BGHIP_EINVAL; }
extern "C" int bghost_model_update_connections(bghost_model *w, bghip_growth_stats *st)
{
    return w ? w->m->updateConnections(st) : BGHIP_EINVAL;
}
extern "C" int bghost_model_update_history(bghost_model *w) { return w ? w->m->updateHistory() : BGHIP_EINVAL; }
extern "C" int bghost_model_run_epoch(bghost_model *w)
{
    if (!w) return BGHIP_EINVAL;
    int rc = w->m->advanceUntilGrowth();
    if (!rc) rc = w->m->updateConnections(nullptr);
    if (!rc) rc = w->m->updateHistory();
    return rc;
}
extern "C" int bghost_model_save_xml(bghost_model *w, const char *path)
{
    return w && path ? w->m->saveData(path) : BGHIP_EINVAL;
}
extern "C" int bghost_model_serialize(bghost_model *w, const char *path)
{
    return w && path ? w->m->serialize(path) : BGHIP_EINVAL;
}
extern "C" int bghost_model_deserialize(bghost_model *w, const char *path)
{
    return w && path ? w->m->deserialize(path) : BGHIP_EINVAL;
}
extern "C" int bghost_model_set_radii(bghost_model *w, float r) { return w ? w->m->setRadii(r) : BGHIP_EINVAL; }
extern "C" const void *bghost_model_array(bghost_model *w, const char *name, int64_t *count)
{
    return w && name ? w->m->array(name, count) : nullptr;
}
extern "C" int32_t bghost_model_epoch(const bghost_model *w) { return w ? w->m->epoch() : -1; }
