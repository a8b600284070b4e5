// vr_runtime.hip - C ABI of the render path (include/vr_api.h): context,
// device buffers, commits, light upload, per-frame launches and read-backs.
// Replaces the reference's OpenCL runtime (template/template.cpp:1266-1754)
// and the device half of World::Render / World::Commit
// (template/world.cpp:2127-2483).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vr_api.h"
#include "vr_kernels.h"
#include "vr_dynamics.h"
#include "vr_mlp.h"

using namespace vr;

static constexpr int kRing = 64;  // frames whose stage events are kept

// contexts alive in the process: a context overlaps its own frames (below)
// only while it is the only one.  With several, their streams already
// overlap each other's work, and a second stream each would crowd the
// process's few hardware queues (measured: 16 Worlds rendered view after
// view, 6.7 -> 8.3 ms per round).
static std::atomic<int> g_liveContexts{0};

// Scratch of vr_trace_rays_noalloc: VR_NOALLOC_MAX_RAYS rays and their hits
// staged on the device (origins and directions 12 B each, voxel, distance and
// side 4 B each: 36 B per ray, 9 MiB).  Every device the library runs on
// holds its own copy from module load until the process ends.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (scratch_address), never through the array's name.
static __device__ __align__(256) unsigned char scratch[(size_t)VR_NOALLOC_MAX_RAYS * 36];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratchMu;
static void* g_scratchPtr[kMaxDevices] = {};

// the device address of `scratch` on `device`, which must be current; looked
// up once per device.  nullptr (and *err set) when the lookup fails.
static void* scratch_address(int device, hipError_t* err)
{
    *err = hipSuccess;
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess) return nullptr;
        g_scratchPtr[device] = p;
    }
    return g_scratchPtr[device];
}

struct vr_ctx {
    vr_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    // this device's `scratch` (vr_trace_rays_noalloc), set by vr_create;
    // nullptr if its lookup failed (noScratch holds why)
    void* scratch = nullptr;
    hipError_t noScratch = hipSuccess;
    // The next frame's primary pass may run on `side` beside this frame's
    // spatial reuse + shading: it writes the hit buffer that only this
    // frame's candidate pass read (`initDone`), unless work queued since then
    // changed the world it traces (`worldEpoch`, advanced by vr_scene_step;
    // vr_commit returns only once its uploads are done).  `side` is created
    // on first use.
    hipStream_t side = nullptr;
    bool counted = false;  // included in g_liveContexts
    hipEvent_t initDone = nullptr, primDone = nullptr;
    // The learned proposal follows the primary pass on `side` too; it reads
    // the network's weights as of their last update, recorded here on
    // `stream` after every training step (other weight and cluster-table
    // changes return only once done).
    hipEvent_t weightsReady = nullptr;
    bool initRecorded = false;
    uint64_t worldEpoch = 0, epochAtInit = 0;
    // start and end events of each stage of the last kRing frames
    hipEvent_t ev[kRing][2 * VR_NUM_STAGES] = {};
    bool pending[kRing] = {};
    uint64_t frameNo = 0;
    int lastSlot = -1;
    bool profiling = false;
    double profSum[VR_NUM_STAGES] = {};
    uint32_t profFrames = 0;
    // world
    uint32_t* grid = nullptr;       // [128^3]
    uint32_t* uber = nullptr;       // [32^3 bits]
    uint32_t* bricks = nullptr;     // [brick_capacity * 512]
    uint32_t* stageIds = nullptr;   // [VR_MAXCOMMITS]
    uint32_t* stageData = nullptr;  // [VR_MAXCOMMITS * 512]
    // lights
    float4* la = nullptr;
    float4* lb = nullptr;
    float4* pick = nullptr;  // power source's alias table (LightView::pick)
    uint32_t nLights = 0, lightCap = 0;
    float* emitLut = nullptr;
    // per-pixel planes
    uint32_t* hitw[2] = {};
    float* dist[2] = {};
    uint32_t* seed[2] = {};
    // reservoir records (ResPlanes), [0] = A (final), [1] = B (initial +
    // temporal): 16 + 16 + 8 bytes per pixel
    uint4* rr0[2] = {};
    uint4* rr1[2] = {};
    float2* rr2[2] = {};
    unsigned long long* rays = nullptr;  // [kRayKinds] traced-ray counters
    bool countRays = false;
    float4* frame = nullptr;
    int cur = 1;              // current hit buffer (swapped before each frame)
    bool committed = false;
    bool framed = false;
    std::string err;
    // learned proposal (allocated on first use)
    bool mlpReady = false;
    float* mlpP = nullptr;       // FP32 master parameters
    float* mlpM = nullptr;       // Adam first moment
    float* mlpV = nullptr;       // Adam second moment
    float* mlpG = nullptr;       // gradient
    __bf16* mlpW1b = nullptr;
    __bf16* mlpW2b = nullptr;
    __bf16* mlpW2Tb = nullptr;
    __bf16* mlpW3b = nullptr;
    __bf16* mlpW3Tb = nullptr;
    uint32_t adamStep = 0;
    uint32_t* clusterStart = nullptr;   // [kMaxClusters + 1]
    uint32_t* clusterOf = nullptr;      // [lights]
    uint4* clusterPick = nullptr;       // [lights] FrameArgs::clusterPick
    float2* clusterPickShare = nullptr; // [lights] FrameArgs::clusterPickShare
    float* clusterPower = nullptr;      // [kMaxClusters] P_c
    float powerLow = 0.0f;              // sum of P_c over clusters with (c & 4) == 0
    float mix = 0.25f;                  // defensive mixture weight of the power source
    uint32_t clusterDraw = VR_CLUSTER_DRAW_POWER;  // the light inside a drawn cluster
    std::vector<vr_light> hostLights;   // the uploaded light list
    uint32_t nclusters = 0, clusterLightCap = 0, clusteredLights = 0;
    uint2* candSlot = nullptr;  // [kCandidates][W*H]: cluster, mass bits
    float* trainX = nullptr;      // training batch
    uint32_t* trainT = nullptr;
    float* trainW = nullptr;
    float* partial = nullptr;     // per-chunk gradients
    float* chunkLoss = nullptr;
    uint32_t trainCap = 0;
    Camera lastCam{};             // camera of the last rendered frame
    uint32_t stageWaves[VR_NUM_STAGES] = {};  // occupancy target per stage (0: default)
    uint32_t stageDda[VR_NUM_STAGES] = {};    // VR_DDA_* of each stage's shadow rays
    // the MIS merge's visibility-ray list and its per-pixel state (allocated
    // by the first MIS frame, the list regrown for more taps)
    uint2* misList = nullptr;
    size_t misCap = 0;             // rays the list holds
    uint32_t* misCount = nullptr;  // [1]
    uint32_t* misOcc = nullptr;    // [W*H]
    uint4* misInfo = nullptr;      // [W*H]
    float2* misTerms = nullptr;    // [taps][W*H], with the list
    // scene motion on the device (allocated by vr_set_light_motion)
    float4* lightPos = nullptr;
    float4* lightVel = nullptr;
    uint32_t* lightVoxel = nullptr;
    uint32_t* lightTarget = nullptr;
    uint32_t* claim = nullptr;    // [kClaimBuckets]
    uint32_t motionCap = 0, motionN = 0;
};

namespace {

int fail(vr_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

inline float bits_float(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

int hipfail(vr_ctx* c, hipError_t e, const char* what)
{
    return fail(c, VR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, call)                                   \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hipfail(ctx, e_, #call); \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count)
{
    return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T) + 16);
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

// Light record: centre = coords + 0.5*size (cl/tools.cl:606-607);
// Le = ToFloatRGB(v) * EmitStrength(v) [* 1/pi] (cl/phat.cl:8-12).
void light_record(const vr_config& cfg, const vr_light& l, float4& a, float4& b)
{
    const uint32_t y = l.position / (1024u * 1024u), z = (l.position / 1024u) % 1024u, x = l.position % 1024u;
    const float hs = 0.5f * (float)l.size;
    a = make_float4((float)x + hs, (float)y + hs, (float)z + hs, (float)l.size);
    const float es = cfg.emit_lut[(l.voxel >> 16) & 255u];
    float r = (float)((l.voxel >> 8) & 15u) * (1.0f / 15.0f);
    float g = (float)((l.voxel >> 4) & 15u) * (1.0f / 15.0f);
    float bl = (float)(l.voxel & 15u) * (1.0f / 15.0f);
    r = r * es, g = g * es, bl = bl * es;
    if (cfg.gfxexp) r = r * 0.31830988618379067153777f, g = g * 0.31830988618379067153777f,
                    bl = bl * 0.31830988618379067153777f;
    uint32_t packed = (y << 20) | (z << 10) | x;
    float pf;
    std::memcpy(&pf, &packed, 4);
    b = make_float4(r, g, bl, pf);
}

// power of a light: rgb sum x emission x emitting area (binary64), the
// weight of the power-proportional source
double light_power(const vr_config& cfg, const vr_light& l)
{
    const double pi = 3.14159265358979323846;
    const double r = (double)((float)((l.voxel >> 8) & 15u) * (1.0f / 15.0f));
    const double g = (double)((float)((l.voxel >> 4) & 15u) * (1.0f / 15.0f));
    const double b = (double)((float)(l.voxel & 15u) * (1.0f / 15.0f));
    const double es = (double)cfg.emit_lut[(l.voxel >> 16) & 255u];
    const double s = (double)l.size;
    const double geo = cfg.light_mode == VR_LIGHTS_POINT ? 4.0 * pi : 6.0 * s * s * pi;
    return (r + g + b) * es * geo;
}

// Vose alias table over n positive weights (binary64): bucket k keeps
// itself with probability prob[k], else alias[k]
void vose(const double* w, uint32_t n, float* prob, uint32_t* alias)
{
    double total = 0;
    for (uint32_t i = 0; i < n; i++) total += w[i];
    std::vector<double> sc(n);
    std::vector<uint32_t> small, large;
    for (uint32_t i = 0; i < n; i++) {
        sc[i] = w[i] / total * (double)n;
        (sc[i] < 1.0 ? small : large).push_back(i);
        prob[i] = 1.0f, alias[i] = i;
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t s = small.back(), l = large.back();
        small.pop_back(), large.pop_back();
        prob[s] = (float)sc[s];
        alias[s] = l;
        sc[l] = (sc[l] + sc[s]) - 1.0;
        (sc[l] < 1.0 ? small : large).push_back(l);
    }
}

// Power-proportional source: Vose alias table over lightContribution
// (cl/tools.cl:579-591) reduced to the rgb sum; binary64 arithmetic.
void power_table(const vr_config& cfg, const vr_light* L, uint32_t n, std::vector<float>& prob,
                 std::vector<uint32_t>& alias, std::vector<float>& pmf)
{
    prob.assign(n, 1.0f), alias.resize(n), pmf.resize(n);
    std::vector<double> pw(n), sc(n);
    std::vector<uint32_t> small, large;
    small.reserve(n), large.reserve(n);
    const double pi = 3.14159265358979323846;
    double total = 0;
    for (uint32_t i = 0; i < n; i++) {
        const double r = (double)((float)((L[i].voxel >> 8) & 15u) * (1.0f / 15.0f));
        const double g = (double)((float)((L[i].voxel >> 4) & 15u) * (1.0f / 15.0f));
        const double b = (double)((float)(L[i].voxel & 15u) * (1.0f / 15.0f));
        const double es = (double)cfg.emit_lut[(L[i].voxel >> 16) & 255u];
        const double s = (double)L[i].size;
        const double geo = cfg.light_mode == VR_LIGHTS_POINT ? 4.0 * pi : 6.0 * s * s * pi;
        pw[i] = (r + g + b) * es * geo;
        total += pw[i];
    }
    if (!(total > 0)) {
        for (uint32_t i = 0; i < n; i++) pw[i] = 1.0;
        total = (double)n;
    }
    for (uint32_t i = 0; i < n; i++) {
        pmf[i] = (float)(pw[i] / total);
        sc[i] = pw[i] / total * (double)n;
        if (sc[i] < 1.0) small.push_back(i);
        else large.push_back(i);
        alias[i] = i;
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t s = small.back(), l = large.back();
        small.pop_back(), large.pop_back();
        prob[s] = (float)sc[s];
        alias[s] = l;
        sc[l] = (sc[l] + sc[s]) - 1.0;
        if (sc[l] < 1.0) small.push_back(l);
        else large.push_back(l);
    }
    for (uint32_t l : large) prob[l] = 1.0f, alias[l] = l;
    for (uint32_t s : small) prob[s] = 1.0f, alias[s] = s;
}

// add a profiled frame's stage times to the running sums (waits for it)
int harvest(vr_ctx* c, int slot)
{
    if (!c->pending[slot]) return VR_OK;
    c->pending[slot] = false;
    HIPCHK(c, hipEventSynchronize(c->ev[slot][2 * VR_NUM_STAGES - 1]));
    for (int s = 0; s < VR_NUM_STAGES; s++) {
        float ms = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev[slot][2 * s], c->ev[slot][2 * s + 1]));
        c->profSum[s] += ms;
    }
    c->profFrames++;
    return VR_OK;
}

// non-finite camera data would make every DDA walk its whole step budget
bool finite_camera(const vr_render_params* p)
{
    const float* vals[] = {p->oneOverRes, p->E, p->p0, p->p1, p->p2, p->Nleft, p->Nright, p->Ntop, p->Nbottom};
    const int counts[] = {2, 3, 3, 3, 3, 4, 4, 4, 4};
    for (int k = 0; k < 9; k++)
        for (int i = 0; i < counts[k]; i++)
            if (!std::isfinite(vals[k][i])) return false;
    return true;
}

FrameArgs frame_args(const vr_ctx* c, const vr_render_params* p)
{
    FrameArgs a;
    std::memset(&a, 0, sizeof(a));
    a.w = WorldView{c->grid, c->bricks, c->uber};
    a.L = LightView{c->la, c->lb, c->pick, p ? p->numberOfLights : 0u};
    if (p) {
        a.cam.ox = p->oneOverRes[0], a.cam.oy = p->oneOverRes[1];
        a.cam.E = f3{p->E[0], p->E[1], p->E[2]};
        a.cam.p0 = f3{p->p0[0], p->p0[1], p->p0[2]};
        a.cam.d10 = f3{p->p1[0] - p->p0[0], p->p1[1] - p->p0[1], p->p1[2] - p->p0[2]};
        a.cam.d20 = f3{p->p2[0] - p->p0[0], p->p2[1] - p->p0[1], p->p2[2] - p->p0[2]};
        a.R0 = p->R0;
        a.restirtemporalframe = p->restirtemporalframe;
        a.temporal = p->temporal ? 1u : 0u;
        a.spatial = p->spatial ? 1u : 0u;
        a.nCand = p->numberOfCandidates;
        a.maxTemporal = p->numberOfMaxTemporalImportance;
        a.spatialRadius = p->spatialRadius;
        a.spatialTaps = p->spatialTaps;
        a.Nl = make_float4(p->Nleft[0], p->Nleft[1], p->Nleft[2], p->Nleft[3]);
        a.Nr = make_float4(p->Nright[0], p->Nright[1], p->Nright[2], p->Nright[3]);
        a.Nt = make_float4(p->Ntop[0], p->Ntop[1], p->Ntop[2], p->Ntop[3]);
        a.Nb = make_float4(p->Nbottom[0], p->Nbottom[1], p->Nbottom[2], p->Nbottom[3]);
    }
    const int cur = c->cur;
    a.cur = HitPlanes{c->hitw[cur], c->dist[cur], c->seed[cur]};
    a.A = ResPlanes{c->rr0[0], c->rr1[0], c->rr2[0]};
    a.B = ResPlanes{c->rr0[1], c->rr1[1], c->rr2[1]};
    a.rays = c->countRays ? c->rays : nullptr;
    a.frame = c->frame;
    a.emitLut = c->emitLut;
    a.W = c->cfg.width, a.H = c->cfg.height;
    a.point = c->cfg.light_mode == VR_LIGHTS_POINT ? 1u : 0u;
    a.gfxexp = c->cfg.gfxexp ? 1u : 0u;
    a.source = c->cfg.source_pdf;
    a.axisOnlyFace = (c->cfg.quirks & VR_QUIRK_AXIS_ONLY_FACE) ? 1u : 0u;
    a.doubleRadiance = (c->cfg.quirks & VR_QUIRK_DOUBLE_RADIANCE) ? 1u : 0u;
    a.shadowSteps = (c->cfg.quirks & VR_QUIRK_SHADOW_STEP_CAP) ? 128 : 999999;
    a.tileOrder = c->cfg.flags & 3u;
    a.misSpatial = (c->cfg.flags & VR_FLAG_MIS_SPATIAL) ? 1u : 0u;
    a.candSlot = c->candSlot;
    a.clusterStart = c->clusterStart;
    a.clusterPick = c->clusterPick;
    a.clusterPickShare = c->clusterPickShare;
    return a;
}

MlpView mlp_view(const vr_ctx* c)
{
    return MlpView{c->mlpW1b, c->mlpW2b, c->mlpW2Tb, c->mlpW3b, c->mlpW3Tb, c->mlpP + kOffB1, c->mlpP + kOffB2,
                   c->mlpP + kOffB3, c->clusterStart, c->clusterPower, c->nclusters, c->mix, c->powerLow};
}

// allocate the network, optimiser state, clusters and candidate planes
int mlp_alloc(vr_ctx* c)
{
    if (c->mlpReady) return VR_OK;
    const size_t npx = (size_t)c->cfg.width * c->cfg.height;
    auto ok = [](hipError_t e) { return e == hipSuccess; };
    const bool good = ok(dalloc(&c->mlpP, kMlpParams)) && ok(dalloc(&c->mlpM, kMlpParams)) &&
                      ok(dalloc(&c->mlpV, kMlpParams)) && ok(dalloc(&c->mlpG, kMlpParams)) &&
                      ok(dalloc(&c->mlpW1b, 64 * 16)) && ok(dalloc(&c->mlpW2b, 64 * 64)) &&
                      ok(dalloc(&c->mlpW2Tb, 64 * 64)) && ok(dalloc(&c->mlpW3b, 256 * 64)) &&
                      ok(dalloc(&c->mlpW3Tb, 256 * 64)) && ok(dalloc(&c->clusterStart, kMaxClusters + 1)) &&
                      ok(dalloc(&c->clusterPower, kMaxClusters)) &&
                      ok(dalloc(&c->candSlot, npx * kCandidates));
    if (!good) {
        // free what was allocated, so that a later call starts clean
        for (void** q : {(void**)&c->mlpP, (void**)&c->mlpM, (void**)&c->mlpV, (void**)&c->mlpG, (void**)&c->mlpW1b,
                         (void**)&c->mlpW2b, (void**)&c->mlpW2Tb, (void**)&c->mlpW3b, (void**)&c->mlpW3Tb,
                         (void**)&c->clusterStart, (void**)&c->clusterPower, (void**)&c->candSlot})
            dfree(*q), *q = nullptr;
        return fail(c, VR_ERR_NOMEM, "learned proposal: device allocation failed");
    }
    // He-uniform weights (bound sqrt(6/fan_in)), zero biases; the output
    // layer is scaled by 1/8 so the first proposals are near-uniform
    std::vector<float> init(kMlpParams, 0.0f);
    uint32_t rs = 1u;
    auto uni = [&rs]() {
        rs ^= rs << 13, rs ^= rs >> 17, rs ^= rs << 5;
        return (float)((double)rs * (2.0 / 4294967296.0) - 1.0);
    };
    auto fill = [&](int off, int n, int fanIn, float scale) {
        const float bound = scale * std::sqrt(6.0f / (float)fanIn);
        for (int i = 0; i < n; i++) init[off + i] = bound * uni();
    };
    fill(kOffW1, 64 * 9, 9, 1.0f);
    fill(kOffW2, 64 * 64, 64, 1.0f);
    fill(kOffW3, 256 * 64, 64, 0.125f);
    HIPCHK(c, hipMemcpyAsync(c->mlpP, init.data(), kMlpParams * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->mlpM, 0, kMlpParams * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->mlpV, 0, kMlpParams * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->clusterStart, 0, (kMaxClusters + 1) * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->clusterPower, 0, kMaxClusters * 4, c->stream));
    HIPCHK(c, launch_mlp_cast(c->mlpP, MlpBuffers{c->mlpW1b, c->mlpW2b, c->mlpW2Tb, c->mlpW3b, c->mlpW3Tb}, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->mlpReady = true;
    return VR_OK;
}

int train_alloc(vr_ctx* c, uint32_t n)
{
    if (n <= c->trainCap) return VR_OK;
    dfree(c->trainX), dfree(c->trainT), dfree(c->trainW), dfree(c->partial), dfree(c->chunkLoss);
    c->trainX = nullptr, c->trainT = nullptr, c->trainW = nullptr, c->partial = nullptr, c->chunkLoss = nullptr;
    c->trainCap = 0;
    const uint32_t chunks = (n + 63) / 64;
    if (dalloc(&c->trainX, (size_t)n * 9) != hipSuccess || dalloc(&c->trainT, n) != hipSuccess ||
        dalloc(&c->trainW, n) != hipSuccess || dalloc(&c->partial, (size_t)chunks * kMlpParams) != hipSuccess ||
        dalloc(&c->chunkLoss, chunks) != hipSuccess)
        return fail(c, VR_ERR_NOMEM, "learned proposal: training buffers");
    c->trainCap = n;
    return VR_OK;
}

// one Adam step from samples [first, first + n) of trainX/T/W; returns the
// batch loss before the update
int train_step(vr_ctx* c, uint32_t first, uint32_t n, float lr, double* loss)
{
    const uint32_t chunks = (n + 63) / 64;
    c->adamStep++;
    HIPCHK(c, launch_train_step(mlp_view(c), c->trainX + (size_t)first * 9, c->trainT + first, c->trainW + first, n,
                                c->partial, c->chunkLoss, c->mlpG, c->mlpP, c->mlpM, c->mlpV,
                                MlpBuffers{c->mlpW1b, c->mlpW2b, c->mlpW2Tb, c->mlpW3b, c->mlpW3Tb}, lr, c->adamStep,
                                c->stream));
    HIPCHK(c, hipEventRecord(c->weightsReady, c->stream));
    if (loss) {
        std::vector<float> l(chunks);
        HIPCHK(c, hipMemcpyAsync(l.data(), c->chunkLoss, (size_t)chunks * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double s = 0;
        for (float v : l) s += v;
        *loss = s;
    }
    return VR_OK;
}

}  // namespace

extern "C" {

void vr_default_config(vr_config* cfg, uint32_t width, uint32_t height)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->width = width, cfg->height = height;
    cfg->light_mode = VR_LIGHTS_AREA;
    cfg->gfxexp = 1;
    cfg->quirks = VR_QUIRKS_REFERENCE;
    cfg->source_pdf = VR_SOURCE_UNIFORM;
    cfg->brick_capacity = VR_BRICKCOUNT;
    for (int i = 0; i < 256; i++) cfg->emit_lut[i] = (float)i;
}

int vr_create(const vr_config* cfg, vr_ctx** out)
{
    if (!cfg || !out) return VR_ERR_ARG;
    *out = nullptr;
    if (cfg->width == 0 || cfg->height == 0 || cfg->width > 16384 || cfg->height > 16384) return VR_ERR_ARG;
    if (cfg->brick_capacity == 0 || cfg->brick_capacity > (1u << 22)) return VR_ERR_ARG;
    if (cfg->source_pdf > VR_SOURCE_LEARNED || cfg->light_mode > VR_LIGHTS_POINT) return VR_ERR_ARG;
    if (cfg->emit_lut[0] != 0.0f) return VR_ERR_ARG;
    if (cfg->flags & ~(((1u << (2 * VR_NUM_STAGES)) - 1u) | VR_FLAG_MIS_SPATIAL)) return VR_ERR_ARG;
    for (int s = 0; s < VR_NUM_STAGES; s++)
        if (((cfg->flags >> (2 * s)) & 3u) > VR_TILES_XCD_BLOCKS) return VR_ERR_ARG;
    vr_ctx* c = new (std::nothrow) vr_ctx();
    if (!c) return VR_ERR_NOMEM;
    c->cfg = *cfg;
    const size_t n = (size_t)cfg->width * cfg->height;
    auto ok = [&](hipError_t e) { return e == hipSuccess; };
    bool good = ok(hipGetDevice(&c->device)) && ok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) &&
                ok(hipEventCreateWithFlags(&c->initDone, hipEventDisableTiming)) &&
                ok(hipEventCreateWithFlags(&c->primDone, hipEventDisableTiming)) &&
                ok(hipEventCreateWithFlags(&c->weightsReady, hipEventDisableTiming));
    // recorded once up front, so that waiting on it before any training
    // step is defined
    good = good && ok(hipEventRecord(c->weightsReady, c->stream));
    for (int f = 0; good && f < kRing; f++)
        for (int i = 0; good && i < 2 * VR_NUM_STAGES; i++) good = ok(hipEventCreate(&c->ev[f][i]));
    good = good && ok(dalloc(&c->grid, VR_GRIDSIZE)) && ok(dalloc(&c->uber, 1024)) &&
           ok(dalloc(&c->bricks, (size_t)cfg->brick_capacity * VR_BRICKSIZE)) &&
           ok(dalloc(&c->stageIds, VR_MAXCOMMITS)) && ok(dalloc(&c->stageData, (size_t)VR_MAXCOMMITS * VR_BRICKSIZE)) &&
           ok(dalloc(&c->emitLut, 256)) && ok(dalloc(&c->frame, n)) && ok(dalloc(&c->rays, (size_t)kRayKinds));
    for (int i = 0; good && i < 2; i++)
        good = ok(dalloc(&c->hitw[i], n)) && ok(dalloc(&c->dist[i], n)) && ok(dalloc(&c->seed[i], n)) &&
               ok(dalloc(&c->rr0[i], n)) && ok(dalloc(&c->rr1[i], n)) && ok(dalloc(&c->rr2[i], n));
    if (good) {
        good = ok(hipMemsetAsync(c->grid, 0, (size_t)VR_GRIDSIZE * 4, c->stream)) &&
               ok(hipMemsetAsync(c->uber, 0, 4096, c->stream)) &&
               ok(hipMemsetAsync(c->bricks, 0, (size_t)cfg->brick_capacity * VR_BRICKSIZE * 4, c->stream)) &&
               ok(hipMemcpyAsync(c->emitLut, cfg->emit_lut, 256 * 4, hipMemcpyHostToDevice, c->stream)) &&
               ok(hipMemsetAsync(c->frame, 0, n * 16, c->stream));
        for (int i = 0; good && i < 2; i++)
            good = ok(hipMemsetAsync(c->hitw[i], 0, n * 4, c->stream)) && ok(hipMemsetAsync(c->dist[i], 0, n * 4, c->stream)) &&
                   ok(hipMemsetAsync(c->seed[i], 0, n * 4, c->stream)) && ok(hipMemsetAsync(c->rr0[i], 0, n * 16, c->stream)) &&
                   ok(hipMemsetAsync(c->rr1[i], 0, n * 16, c->stream)) && ok(hipMemsetAsync(c->rr2[i], 0, n * 8, c->stream));
        good = good && ok(hipMemsetAsync(c->rays, 0, kRayKinds * 8, c->stream));
        good = good && ok(hipStreamSynchronize(c->stream));
    }
    if (!good) {
        vr_destroy(c);
        return VR_ERR_NOMEM;
    }
    // set-up of the allocation-free entry point, with the context's device
    // current: a failed lookup leaves every other entry point working
    c->scratch = scratch_address(c->device, &c->noScratch);
    g_liveContexts++;
    c->counted = true;
    *out = c;
    return VR_OK;
}

int vr_destroy(vr_ctx* c)
{
    if (!c) return VR_ERR_ARG;
    if (c->counted) g_liveContexts--;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    dfree(c->grid), dfree(c->uber), dfree(c->bricks), dfree(c->stageIds), dfree(c->stageData);
    dfree(c->la), dfree(c->lb), dfree(c->pick), dfree(c->emitLut), dfree(c->frame);
    dfree(c->rays);
    for (int i = 0; i < 2; i++) {
        dfree(c->hitw[i]), dfree(c->dist[i]), dfree(c->seed[i]);
        dfree(c->rr0[i]), dfree(c->rr1[i]), dfree(c->rr2[i]);
    }
    dfree(c->mlpP), dfree(c->mlpM), dfree(c->mlpV), dfree(c->mlpG);
    dfree(c->mlpW1b), dfree(c->mlpW2b), dfree(c->mlpW2Tb), dfree(c->mlpW3b), dfree(c->mlpW3Tb);
    dfree(c->clusterStart), dfree(c->clusterOf), dfree(c->clusterPick), dfree(c->clusterPickShare);
    dfree(c->clusterPower);
    dfree(c->candSlot);
    dfree(c->misList), dfree(c->misCount), dfree(c->misOcc), dfree(c->misInfo), dfree(c->misTerms);
    dfree(c->trainX), dfree(c->trainT), dfree(c->trainW), dfree(c->partial), dfree(c->chunkLoss);
    dfree(c->lightPos), dfree(c->lightVel), dfree(c->lightVoxel), dfree(c->lightTarget), dfree(c->claim);
    for (int i = 0; i < 2 * VR_NUM_STAGES; i++)
        for (int f = 0; f < kRing; f++)
            if (c->ev[f][i]) (void)hipEventDestroy(c->ev[f][i]);
    if (c->initDone) (void)hipEventDestroy(c->initDone);
    if (c->primDone) (void)hipEventDestroy(c->primDone);
    if (c->weightsReady) (void)hipEventDestroy(c->weightsReady);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VR_OK;
}

const char* vr_last_error(const vr_ctx* c) { return c ? c->err.c_str() : "null context"; }

int vr_commit(vr_ctx* c, const uint32_t* grid, const uint32_t* ids, const uint32_t* payload, uint32_t n)
{
    if (!c) return VR_ERR_ARG;
    if (!grid) return fail(c, VR_ERR_ARG, "vr_commit: grid is null");
    if (n > 0 && (!ids || !payload)) return fail(c, VR_ERR_ARG, "vr_commit: brick ids/payload null");
    const uint32_t cap = c->cfg.brick_capacity;
    // every brick cell must name a brick inside the pool: the traversal
    // indexes the pool with it
    for (uint32_t i = 0; i < VR_GRIDSIZE; i++)
        if ((grid[i] & 1u) && (grid[i] >> 1) >= cap)
            return fail(c, VR_ERR_ARG, "vr_commit: grid cell " + std::to_string(i) + " names brick " +
                                           std::to_string(grid[i] >> 1) + " beyond capacity " + std::to_string(cap));
    for (uint32_t i = 0; i < n; i++)
        if (ids[i] >= cap)
            return fail(c, VR_ERR_ARG, "vr_commit: brick id " + std::to_string(ids[i]) + " beyond capacity");
    HIPCHK(c, hipMemcpyAsync(c->grid, grid, (size_t)VR_GRIDSIZE * 4, hipMemcpyHostToDevice, c->stream));
    for (uint32_t done = 0; done < n;) {
        const uint32_t k = (n - done) < VR_MAXCOMMITS ? (n - done) : VR_MAXCOMMITS;
        HIPCHK(c, hipMemcpyAsync(c->stageIds, ids + done, (size_t)k * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->stageData, payload + (size_t)done * VR_BRICKSIZE, (size_t)k * VR_BRICKSIZE * 4,
                                 hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launch_commit(c->stageIds, c->stageData, c->bricks, k, cap, c->stream));
        done += k;
    }
    HIPCHK(c, launch_uber(c->grid, c->uber, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->committed = true;
    return VR_OK;
}

int vr_set_lights(vr_ctx* c, const vr_light* lights, uint32_t n)
{
    if (!c) return VR_ERR_ARG;
    if (n > 0 && !lights) return fail(c, VR_ERR_ARG, "vr_set_lights: null lights");
    std::vector<float4> ha(n), hb(n);
    for (uint32_t i = 0; i < n; i++) {
        if (lights[i].position >= (1u << 30) || lights[i].size == 0 || lights[i].size > 1024)
            return fail(c, VR_ERR_ARG, "vr_set_lights: light " + std::to_string(i) + " out of range");
        light_record(c->cfg, lights[i], ha[i], hb[i]);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n > c->lightCap) {
        dfree(c->la), dfree(c->lb), dfree(c->pick);
        c->la = c->lb = c->pick = nullptr;
        const uint32_t cap = n * 2;
        if (dalloc(&c->la, cap) != hipSuccess || dalloc(&c->lb, cap) != hipSuccess ||
            dalloc(&c->pick, cap) != hipSuccess) {
            // the old list's buffers are gone: leave an empty list, so that
            // no later frame or step can reach the freed records
            c->lightCap = 0, c->nLights = 0, c->motionN = 0;
            c->hostLights.clear();
            return fail(c, VR_ERR_NOMEM, "vr_set_lights: device allocation failed");
        }
        c->lightCap = cap;
    }
    c->nLights = n;
    c->hostLights.assign(lights, lights + n);
    c->motionN = 0;  // a new list has no motion until vr_set_light_motion
    // the cluster tables describe the previous list, even at the same size
    c->clusteredLights = 0;
    if (n == 0) return VR_OK;
    HIPCHK(c, hipMemcpyAsync(c->la, ha.data(), (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->lb, hb.data(), (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    if (c->cfg.source_pdf == VR_SOURCE_POWER) {
        std::vector<float> prob, pmf;
        std::vector<uint32_t> alias;
        power_table(c->cfg, lights, n, prob, alias, pmf);
        std::vector<float4> pick(n);
        for (uint32_t i = 0; i < n; i++) {
            uint32_t a = alias[i];
            float ab;
            std::memcpy(&ab, &a, 4);
            pick[i] = make_float4(prob[i], ab, pmf[i], pmf[a]);
        }
        HIPCHK(c, hipMemcpyAsync(c->pick, pick.data(), (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VR_OK;
}

// the checks of vr_render_frame (also applied to every view of a batch)
static int check_frame(vr_ctx* c, const vr_render_params* p, const char* who)
{
    if (!c->committed) return fail(c, VR_ERR_STATE, std::string(who) + ": no world committed");
    if (p->skyDomeSampling) return fail(c, VR_ERR_UNSUPPORTED, std::string(who) + ": sky dome sampling is not supported");
    if (p->numberOfLights != c->nLights)
        return fail(c, VR_ERR_ARG, std::string(who) + ": numberOfLights " + std::to_string(p->numberOfLights) +
                                       " differs from the uploaded light count " + std::to_string(c->nLights));
    if (p->spatialRadius == 0 && p->spatial && p->spatialTaps)
        return fail(c, VR_ERR_ARG, std::string(who) + ": spatialRadius must be >= 1");
    if (!finite_camera(p)) return fail(c, VR_ERR_ARG, std::string(who) + ": non-finite camera parameters");
    if ((c->cfg.flags & VR_FLAG_MIS_SPATIAL) && p->spatial && p->spatialTaps > VR_MIS_MAX_TAPS)
        return fail(c, VR_ERR_ARG, std::string(who) + ": the MIS spatial merge takes at most " +
                                       std::to_string(VR_MIS_MAX_TAPS) + " taps");
    if (c->cfg.source_pdf == VR_SOURCE_LEARNED) {
        if (!c->mlpReady || c->nclusters == 0 || c->clusteredLights != c->nLights)
            return fail(c, VR_ERR_STATE, std::string(who) + ": the learned source needs vr_set_light_clusters for the "
                                         "current light list");
        if (p->numberOfCandidates > (uint32_t)kCandidates)
            return fail(c, VR_ERR_ARG, std::string(who) + ": the learned source draws at most 32 candidates");
    }
    return VR_OK;
}

// The MIS merge's ray list for `taps` taps per pixel.  Pixel indices share
// a list word with the tap index (5 bits), so frames of more than 2^27
// pixels keep the one-kernel merge (mis_split false).  Regrowing waits for
// the frames in flight, which may still read the old list.
static int mis_alloc(vr_ctx* c, uint32_t taps, bool& mis_split)
{
    const size_t npx = (size_t)c->cfg.width * c->cfg.height;
    mis_split = npx <= ((size_t)1 << 27);
    if (!mis_split || taps == 0u) return VR_OK;
    const size_t need = npx * taps;
    if (c->misList && c->misCap >= need) return VR_OK;
    if (!c->misCount) {
        if (dalloc(&c->misCount, 1) != hipSuccess || dalloc(&c->misOcc, npx) != hipSuccess ||
            dalloc(&c->misInfo, npx) != hipSuccess) {
            dfree(c->misCount), dfree(c->misOcc), dfree(c->misInfo);
            c->misCount = nullptr, c->misOcc = nullptr, c->misInfo = nullptr;
            return fail(c, VR_ERR_NOMEM, "MIS merge: device allocation failed");
        }
        HIPCHK(c, hipMemset(c->misCount, 0, sizeof(uint32_t)));
    }
    if (c->misList) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->side) HIPCHK(c, hipStreamSynchronize(c->side));
        dfree(c->misList), dfree(c->misTerms);
        c->misList = nullptr, c->misTerms = nullptr, c->misCap = 0;
    }
    if (dalloc(&c->misList, need) != hipSuccess || dalloc(&c->misTerms, need) != hipSuccess) {
        dfree(c->misList), dfree(c->misTerms);
        c->misList = nullptr, c->misTerms = nullptr;
        return fail(c, VR_ERR_NOMEM, "MIS merge: device allocation of the ray list failed");
    }
    c->misCap = need;
    return VR_OK;
}

int vr_render_frame(vr_ctx* c, const vr_render_params* p)
{
    if (!c || !p) return VR_ERR_ARG;
    {
        const int rc = check_frame(c, p, "vr_render_frame");
        if (rc != VR_OK) return rc;
    }
    const bool learned = c->cfg.source_pdf == VR_SOURCE_LEARNED;
    bool misSplit = false;
    if ((c->cfg.flags & VR_FLAG_MIS_SPATIAL) && p->spatial) {
        const int rc = mis_alloc(c, p->spatialTaps, misSplit);
        if (rc != VR_OK) return rc;
        misSplit = misSplit && c->misList != nullptr;
    }
    const int slot = (int)(c->frameNo % kRing);
    {
        // before any state changes: a failure here leaves the context as it was
        const int rc = harvest(c, slot);
        if (rc != VR_OK) return rc;
    }
    c->cur = 1 - c->cur;  // primaryHitBuffer swap, world.cpp:2221-2223
    const FrameArgs a = frame_args(c, p);
    // the primary pass beside the previous frame's last stage when nothing
    // since that frame's candidate pass has changed the world
    bool overlap = c->initRecorded && c->epochAtInit == c->worldEpoch && g_liveContexts.load() == 1;
    if (overlap && !c->side && hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
        c->side = nullptr;
        overlap = false;
    }
    hipStream_t ps = overlap ? c->side : c->stream;
    if (overlap) HIPCHK(c, hipStreamWaitEvent(c->side, c->initDone, 0));
    for (int s = 0; s < VR_NUM_STAGES; s++) {
        FrameArgs as = a;
        as.tileOrder = (c->cfg.flags >> (2 * s)) & 3u;
        as.waves = c->stageWaves[s];
        as.dda = c->stageDda[s];
        hipStream_t st = s == VR_STAGE_PRIMARY ? ps : c->stream;
        // the proposal runs after the primary hits and is timed with the
        // candidate stage: beside the previous frame's last stage too when
        // the primary pass is (it writes the slots that only the previous
        // frame's candidate pass read, and waits for the last training step)
        const bool propose = s == VR_STAGE_INITIAL && learned;
        if (s == VR_STAGE_INITIAL && overlap) {
            if (propose) {
                HIPCHK(c, hipStreamWaitEvent(c->side, c->weightsReady, 0));
                HIPCHK(c, hipEventRecord(c->ev[slot][2 * s], c->side));
                HIPCHK(c, launch_proposal(as, mlp_view(c), c->candSlot, proposal_salt(p->R0), c->side));
            }
            HIPCHK(c, hipEventRecord(c->primDone, c->side));
            HIPCHK(c, hipStreamWaitEvent(c->stream, c->primDone, 0));
        }
        if (!(propose && overlap)) HIPCHK(c, hipEventRecord(c->ev[slot][2 * s], st));
        if (propose && !overlap)
            HIPCHK(c, launch_proposal(as, mlp_view(c), c->candSlot, proposal_salt(p->R0), c->stream));
        if (s == VR_STAGE_RESOLVE && misSplit) {
            const MisBufs mb{c->misList, c->misCount, c->misOcc, c->misInfo, c->misTerms};
            HIPCHK(c, launch_mis_split(as, mb, (uint32_t)((c->misCap + 255) / 256), st));
        } else {
            HIPCHK(c, launch_stage(s, as, st));
        }
        HIPCHK(c, hipEventRecord(c->ev[slot][2 * s + 1], st));
        if (s == VR_STAGE_INITIAL) {
            HIPCHK(c, hipEventRecord(c->initDone, c->stream));
            c->initRecorded = true;
            c->epochAtInit = c->worldEpoch;
        }
    }
    c->pending[slot] = c->profiling;
    c->lastSlot = slot;
    c->lastCam = a.cam;
    c->frameNo++;
    c->framed = true;
    return VR_OK;
}

int vr_sync(vr_ctx* c)
{
    if (!c) return VR_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VR_OK;
}

int vr_set_stage_waves(vr_ctx* c, int32_t stage, uint32_t waves)
{
    if (!c || stage < -1 || stage >= VR_NUM_STAGES || (waves != 0 && (waves < 5 || waves > 8))) return VR_ERR_ARG;
    for (int s = 0; s < VR_NUM_STAGES; s++)
        if (stage < 0 || stage == s) c->stageWaves[s] = waves;
    return VR_OK;
}

static_assert(VR_DDA_SCHEDULED == kDdaScheduled && VR_DDA_NESTED == kDdaNested, "DDA mode values");

int vr_set_stage_dda(vr_ctx* c, int32_t stage, uint32_t mode)
{
    if (!c) return VR_ERR_ARG;
    if (mode > VR_DDA_NESTED) return fail(c, VR_ERR_ARG, "vr_set_stage_dda: unknown mode");
    if (stage < -1 || stage == VR_STAGE_PRIMARY || stage >= VR_NUM_STAGES)
        return fail(c, VR_ERR_ARG, "vr_set_stage_dda: unknown stage or the primary stage");
    for (int s = VR_STAGE_INITIAL; s < VR_NUM_STAGES; s++)
        if (stage < 0 || stage == s) c->stageDda[s] = mode;
    return VR_OK;
}

int vr_set_tile_order(vr_ctx* c, int32_t stage, uint32_t order)
{
    if (!c) return VR_ERR_ARG;
    if (order > VR_TILES_XCD_BLOCKS) return fail(c, VR_ERR_ARG, "vr_set_tile_order: unknown order");
    if (stage < -1 || stage >= VR_NUM_STAGES) return fail(c, VR_ERR_ARG, "vr_set_tile_order: unknown stage");
    for (int s = 0; s < VR_NUM_STAGES; s++)
        if (stage < 0 || stage == s) c->cfg.flags = (c->cfg.flags & ~(3u << (2 * s))) | (order << (2 * s));
    return VR_OK;
}

int vr_read_hits(vr_ctx* c, uint32_t* hitw, float* dist, uint32_t* seed)
{
    if (!c) return VR_ERR_ARG;
    const size_t n = (size_t)c->cfg.width * c->cfg.height;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (hitw) HIPCHK(c, hipMemcpy(hitw, c->hitw[c->cur], n * 4, hipMemcpyDeviceToHost));
    if (dist) HIPCHK(c, hipMemcpy(dist, c->dist[c->cur], n * 4, hipMemcpyDeviceToHost));
    if (seed) HIPCHK(c, hipMemcpy(seed, c->seed[c->cur], n * 4, hipMemcpyDeviceToHost));
    return VR_OK;
}

int vr_read_reservoirs(vr_ctx* c, int which, float* a4, float* b4, uint32_t* m)
{
    if (!c || which < 0 || which > 1) return VR_ERR_ARG;
    const size_t n = (size_t)c->cfg.width * c->cfg.height;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the records (ResPlanes) rearranged into the documented planes
    std::vector<uint4> r0(n), r1(n);
    std::vector<float2> r2(n);
    HIPCHK(c, hipMemcpy(r0.data(), c->rr0[which], n * 16, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(r1.data(), c->rr1[which], n * 16, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(r2.data(), c->rr2[which], n * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        if (a4) {
            a4[4 * i + 0] = r2[i].x, a4[4 * i + 1] = bits_float(r0[i].z);
            a4[4 * i + 2] = r2[i].y, a4[4 * i + 3] = bits_float(r0[i].w);
        }
        if (b4) {
            b4[4 * i + 0] = bits_float(r1[i].x), b4[4 * i + 1] = bits_float(r1[i].y);
            b4[4 * i + 2] = bits_float(r1[i].z), b4[4 * i + 3] = bits_float(r1[i].w);
        }
        if (m) m[i] = r0[i].x & 0x07ffffffu;  // M | face << 24
    }
    return VR_OK;
}

int vr_read_frame(vr_ctx* c, float* rgba)
{
    if (!c || !rgba) return VR_ERR_ARG;
    const size_t n = (size_t)c->cfg.width * c->cfg.height;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(rgba, c->frame, n * 16, hipMemcpyDeviceToHost));
    return VR_OK;
}

int vr_set_ray_counting(vr_ctx* c, int enable)
{
    if (!c) return VR_ERR_ARG;
    if (enable) {
        // the queued frames end first: a counted frame's primary pass may
        // run on the side stream, which a zeroing queued on the context's
        // stream would not be ordered against
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->side) HIPCHK(c, hipStreamSynchronize(c->side));
        HIPCHK(c, hipMemsetAsync(c->rays, 0, kRayKinds * 8, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->countRays = enable != 0;
    return VR_OK;
}

int vr_read_ray_counts(vr_ctx* c, uint64_t counts[VR_RAY_KINDS])
{
    if (!c || !counts) return VR_ERR_ARG;
    static_assert(VR_RAY_KINDS == kRayKinds, "ray kinds");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->side) HIPCHK(c, hipStreamSynchronize(c->side));
    unsigned long long v[kRayKinds];
    HIPCHK(c, hipMemcpy(v, c->rays, sizeof v, hipMemcpyDeviceToHost));
    for (int k = 0; k < kRayKinds; k++) counts[k] = v[k];
    return VR_OK;
}

int vr_stage_times(vr_ctx* c, float ms[VR_NUM_STAGES])
{
    if (!c || !ms) return VR_ERR_ARG;
    if (!c->framed) return fail(c, VR_ERR_STATE, "vr_stage_times: no frame rendered");
    if (c->lastSlot < 0)
        return fail(c, VR_ERR_STATE, "vr_stage_times: the last frame was rendered in a batch (vr_batch_profile_read)");
    const int slot = c->lastSlot;
    HIPCHK(c, hipEventSynchronize(c->ev[slot][2 * VR_NUM_STAGES - 1]));
    for (int s = 0; s < VR_NUM_STAGES; s++)
        HIPCHK(c, hipEventElapsedTime(&ms[s], c->ev[slot][2 * s], c->ev[slot][2 * s + 1]));
    return VR_OK;
}

int vr_profile_begin(vr_ctx* c)
{
    if (!c) return VR_ERR_ARG;
    c->profiling = false;
    for (int f = 0; f < kRing; f++) c->pending[f] = false;
    for (int s = 0; s < VR_NUM_STAGES; s++) c->profSum[s] = 0.0;
    c->profFrames = 0;
    c->profiling = true;
    return VR_OK;
}

int vr_profile_read(vr_ctx* c, double sum_ms[VR_NUM_STAGES], uint32_t* frames)
{
    if (!c || !sum_ms || !frames) return VR_ERR_ARG;
    for (int f = 0; f < kRing; f++) {
        const int rc = harvest(c, f);
        if (rc != VR_OK) return rc;
    }
    for (int s = 0; s < VR_NUM_STAGES; s++) sum_ms[s] = c->profSum[s];
    *frames = c->profFrames;
    return VR_OK;
}

int vr_trace_rays(vr_ctx* c, const float* O, const float* D, uint32_t n, uint32_t kind, int32_t steps,
                  uint32_t* voxel, float* dist, uint32_t* side)
{
    if (!c) return VR_ERR_ARG;
    if (n == 0) return VR_OK;
    if (!O || !D || !voxel || !dist || !side || kind > VR_TRACE_OPAQUE || n > (1u << 26))
        return fail(c, VR_ERR_ARG, "vr_trace_rays: bad arguments");
    if (!c->committed) return fail(c, VR_ERR_STATE, "vr_trace_rays: no world committed");
    for (size_t i = 0; i < (size_t)n * 3; i++)
        if (!std::isfinite(O[i]) || !std::isfinite(D[i]))
            return fail(c, VR_ERR_ARG, "vr_trace_rays: non-finite ray " + std::to_string(i / 3));
    float *dO = nullptr, *dD = nullptr, *dd = nullptr;
    uint32_t *dv = nullptr, *ds = nullptr;
    int rc = VR_OK;
    if (dalloc(&dO, (size_t)n * 3) != hipSuccess || dalloc(&dD, (size_t)n * 3) != hipSuccess ||
        dalloc(&dd, n) != hipSuccess || dalloc(&dv, n) != hipSuccess || dalloc(&ds, n) != hipSuccess) {
        rc = fail(c, VR_ERR_NOMEM, "vr_trace_rays: allocation failed");
    } else {
        hipError_t e = hipMemcpyAsync(dO, O, (size_t)n * 12, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dD, D, (size_t)n * 12, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = launch_trace_batch(WorldView{c->grid, c->bricks, c->uber}, dO, dD, n, kind, steps, dv, dd, ds, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(voxel, dv, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dist, dd, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(side, ds, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hipfail(c, e, "vr_trace_rays");
    }
    (void)hipStreamSynchronize(c->stream);
    dfree(dO), dfree(dD), dfree(dd), dfree(dv), dfree(ds);
    return rc;
}

int vr_trace_rays_noalloc(vr_ctx* c, const float* O, const float* D, uint32_t n, uint32_t kind, int32_t steps,
                          uint32_t* voxel, float* dist, uint32_t* side)
{
    // vr_trace_rays with the rays and hits staged in the device's `scratch`
    // instead of buffers allocated per call: no hipMalloc or hipFree, so the
    // call suits callers that cannot allocate.  The kernel is the same one.
    if (!c) return VR_ERR_ARG;
    if (n == 0) return VR_OK;
    if (!O || !D || !voxel || !dist || !side || kind > VR_TRACE_OPAQUE)
        return fail(c, VR_ERR_ARG, "vr_trace_rays_noalloc: bad arguments");
    if (n > VR_NOALLOC_MAX_RAYS)
        return fail(c, VR_ERR_ARG, "vr_trace_rays_noalloc: " + std::to_string(n) + " rays exceed the scratch's " +
                                       std::to_string(VR_NOALLOC_MAX_RAYS));
    if (!c->committed) return fail(c, VR_ERR_STATE, "vr_trace_rays_noalloc: no world committed");
    if (!c->scratch) return hipfail(c, c->noScratch, "vr_trace_rays_noalloc: scratch address");
    for (size_t i = 0; i < (size_t)n * 3; i++)
        if (!std::isfinite(O[i]) || !std::isfinite(D[i]))
            return fail(c, VR_ERR_ARG, "vr_trace_rays_noalloc: non-finite ray " + std::to_string(i / 3));
    // every byte a launch reads is written by this call's copies first, and
    // every output byte by the kernel, so nothing needs clearing
    unsigned char* base = static_cast<unsigned char*>(c->scratch);
    const size_t cap = VR_NOALLOC_MAX_RAYS;
    float* dO = reinterpret_cast<float*>(base);
    float* dD = reinterpret_cast<float*>(base + cap * 12);
    uint32_t* dv = reinterpret_cast<uint32_t*>(base + cap * 24);
    float* dd = reinterpret_cast<float*>(base + cap * 28);
    uint32_t* ds = reinterpret_cast<uint32_t*>(base + cap * 32);
    hipError_t e = hipMemcpyAsync(dO, O, (size_t)n * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dD, D, (size_t)n * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = launch_trace_batch(WorldView{c->grid, c->bricks, c->uber}, dO, dD, n, kind, steps, dv, dd, ds, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(voxel, dv, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dist, dd, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(side, ds, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(c->stream);
        return hipfail(c, e, "vr_trace_rays_noalloc");
    }
    return VR_OK;
}

int vr_set_light_clusters(vr_ctx* c, const uint32_t* clusterOf, uint32_t nlights, uint32_t nclusters)
{
    if (!c || (nlights > 0 && !clusterOf)) return VR_ERR_ARG;
    if (nclusters == 0 || nclusters > (uint32_t)kMaxClusters)
        return fail(c, VR_ERR_ARG, "vr_set_light_clusters: 1..256 clusters");
    if (nlights != c->nLights || c->hostLights.size() != nlights)
        return fail(c, VR_ERR_ARG, "vr_set_light_clusters: cluster list must cover the uploaded lights");
    for (uint32_t i = 0; i < nlights; i++)
        if (clusterOf[i] >= nclusters) return fail(c, VR_ERR_ARG, "vr_set_light_clusters: cluster id out of range");
    const int rc = mlp_alloc(c);
    if (rc != VR_OK) return rc;
    // counting sort of the light indices by cluster
    std::vector<uint32_t> start(kMaxClusters + 1, 0u), lights(nlights);
    for (uint32_t i = 0; i < nlights; i++) start[clusterOf[i] + 1]++;
    for (uint32_t k = 0; k < (uint32_t)kMaxClusters; k++) start[k + 1] += start[k];
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (uint32_t i = 0; i < nlights; i++) lights[fill[clusterOf[i]]++] = i;
    // power of every light (the power source's pmf), each cluster's share
    // P_c, the in-cluster share pmf_i / P_c and a power alias table per cluster
    std::vector<double> pmf(nlights);
    {
        double total = 0;
        for (uint32_t i = 0; i < nlights; i++) total += (pmf[i] = light_power(c->cfg, c->hostLights[i]));
        for (uint32_t i = 0; i < nlights; i++) pmf[i] = total > 0 ? pmf[i] / total : 1.0 / nlights;
    }
    std::vector<float> power(kMaxClusters, 0.0f), share(nlights, 0.0f), prob(nlights, 1.0f);
    std::vector<uint32_t> alias(nlights, 0u);
    double low = 0;
    for (uint32_t k = 0; k < (uint32_t)kMaxClusters; k++) {
        const uint32_t c0 = start[k], cnt = start[k + 1] - c0;
        if (cnt == 0) continue;
        std::vector<double> w(cnt);
        double pc = 0;
        for (uint32_t j = 0; j < cnt; j++) pc += (w[j] = pmf[lights[c0 + j]]);
        if (!(pc > 0) || c->clusterDraw == VR_CLUSTER_DRAW_UNIFORM) std::fill(w.begin(), w.end(), 1.0);
        for (uint32_t j = 0; j < cnt; j++)
            share[lights[c0 + j]] = (float)(pc > 0 && c->clusterDraw == VR_CLUSTER_DRAW_POWER ? w[j] / pc : 1.0 / cnt);
        vose(w.data(), cnt, prob.data() + c0, alias.data() + c0);
        power[k] = (float)pc;
        if ((k & 4u) == 0u) low += pc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // bucket j of cluster k (global index c0 + j): prob, its own light, its
    // alias's light and local index, and both lights' shares
    std::vector<uint4> pick(nlights);
    std::vector<float2> pickShare(nlights);
    for (uint32_t k = 0; k < (uint32_t)kMaxClusters; k++) {
        const uint32_t c0 = start[k], cnt = start[k + 1] - c0;
        for (uint32_t j = 0; j < cnt; j++) {
            const uint32_t self = lights[c0 + j], other = lights[c0 + alias[c0 + j]];
            uint32_t pb;
            std::memcpy(&pb, &prob[c0 + j], 4);
            pick[c0 + j] = make_uint4(pb, self, other, alias[c0 + j]);
            pickShare[c0 + j] = make_float2(share[self], share[other]);
        }
    }
    if (nlights > c->clusterLightCap) {
        dfree(c->clusterOf), dfree(c->clusterPick), dfree(c->clusterPickShare);
        c->clusterOf = nullptr, c->clusterPick = nullptr, c->clusterPickShare = nullptr, c->clusterLightCap = 0;
        // the old tables are gone: no frame may use them if this fails
        c->nclusters = 0, c->clusteredLights = 0;
        if (dalloc(&c->clusterOf, nlights) != hipSuccess || dalloc(&c->clusterPick, nlights) != hipSuccess ||
            dalloc(&c->clusterPickShare, nlights) != hipSuccess)
            return fail(c, VR_ERR_NOMEM, "vr_set_light_clusters: allocation failed");
        c->clusterLightCap = nlights;
    }
    HIPCHK(c, hipMemcpyAsync(c->clusterStart, start.data(), (kMaxClusters + 1) * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->clusterPower, power.data(), kMaxClusters * 4, hipMemcpyHostToDevice, c->stream));
    if (nlights) {
        HIPCHK(c, hipMemcpyAsync(c->clusterOf, clusterOf, (size_t)nlights * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->clusterPick, pick.data(), (size_t)nlights * 16, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->clusterPickShare, pickShare.data(), (size_t)nlights * 8, hipMemcpyHostToDevice,
                                 c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->powerLow = (float)low;
    c->nclusters = nclusters;
    c->clusteredLights = nlights;
    return VR_OK;
}

int vr_set_cluster_draw(vr_ctx* c, uint32_t mode)
{
    if (!c) return VR_ERR_ARG;
    if (mode > VR_CLUSTER_DRAW_UNIFORM) return fail(c, VR_ERR_ARG, "vr_set_cluster_draw: unknown mode");
    c->clusterDraw = mode;
    c->clusteredLights = 0;  // the tables follow at the next vr_set_light_clusters
    return VR_OK;
}

int vr_set_proposal_mix(vr_ctx* c, float eps)
{
    if (!c) return VR_ERR_ARG;
    if (!(eps >= 0.0f && eps <= 1.0f)) return fail(c, VR_ERR_ARG, "vr_set_proposal_mix: eps must lie in [0, 1]");
    c->mix = eps;
    return VR_OK;
}

int vr_read_cluster_tables(vr_ctx* c, uint32_t* start, uint32_t* lights, float* prob, uint32_t* alias, float* share,
                           float* power)
{
    if (!c) return VR_ERR_ARG;
    if (!c->mlpReady || c->nclusters == 0) return fail(c, VR_ERR_STATE, "vr_read_cluster_tables: no clusters");
    const size_t n = c->clusteredLights;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (start) HIPCHK(c, hipMemcpy(start, c->clusterStart, (kMaxClusters + 1) * 4, hipMemcpyDeviceToHost));
    if (power) HIPCHK(c, hipMemcpy(power, c->clusterPower, kMaxClusters * 4, hipMemcpyDeviceToHost));
    if (n == 0) return VR_OK;
    // the tables as the kernels read them, unpacked
    std::vector<uint4> pick(n);
    std::vector<float2> pickShare(n);
    HIPCHK(c, hipMemcpy(pick.data(), c->clusterPick, n * 16, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(pickShare.data(), c->clusterPickShare, n * 8, hipMemcpyDeviceToHost));
    for (size_t j = 0; j < n; j++) {
        if (lights) lights[j] = pick[j].y;
        if (prob) std::memcpy(&prob[j], &pick[j].x, 4);
        if (alias) alias[j] = pick[j].w;
        if (share && pick[j].y < n) share[pick[j].y] = pickShare[j].x;
    }
    return VR_OK;
}

int vr_mlp_set_params(vr_ctx* c, const float* params, uint32_t n)
{
    if (!c || !params || n != (uint32_t)kMlpParams) return VR_ERR_ARG;
    const int rc = mlp_alloc(c);
    if (rc != VR_OK) return rc;
    HIPCHK(c, hipMemcpyAsync(c->mlpP, params, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->mlpM, 0, kMlpParams * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->mlpV, 0, kMlpParams * 4, c->stream));
    c->adamStep = 0;
    HIPCHK(c, launch_mlp_cast(c->mlpP, MlpBuffers{c->mlpW1b, c->mlpW2b, c->mlpW2Tb, c->mlpW3b, c->mlpW3Tb}, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VR_OK;
}

int vr_mlp_get_params(vr_ctx* c, float* params, uint32_t n)
{
    if (!c || !params || n != (uint32_t)kMlpParams) return VR_ERR_ARG;
    if (!c->mlpReady) return fail(c, VR_ERR_STATE, "vr_mlp_get_params: no network");
    HIPCHK(c, hipMemcpyAsync(params, c->mlpP, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VR_OK;
}

int vr_mlp_probs(vr_ctx* c, const float* x, uint32_t B, float* q)
{
    if (!c || !x || !q || B == 0 || B > (1u << 22)) return VR_ERR_ARG;
    if (!c->mlpReady || c->nclusters == 0) return fail(c, VR_ERR_STATE, "vr_mlp_probs: no network or clusters");
    float *dx = nullptr, *dq = nullptr;
    int rc = VR_OK;
    if (dalloc(&dx, (size_t)B * 9) != hipSuccess || dalloc(&dq, (size_t)B * 256) != hipSuccess) {
        rc = fail(c, VR_ERR_NOMEM, "vr_mlp_probs: allocation failed");
    } else {
        hipError_t e = hipMemcpyAsync(dx, x, (size_t)B * 36, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = launch_mlp_probs(mlp_view(c), dx, B, dq, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(q, dq, (size_t)B * 1024, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hipfail(c, e, "vr_mlp_probs");
    }
    (void)hipStreamSynchronize(c->stream);
    dfree(dx), dfree(dq);
    return rc;
}

int vr_mlp_grad(vr_ctx* c, const float* x, const uint32_t* t, const float* w, uint32_t B, float* grad, double* loss)
{
    if (!c || !x || !t || !w || !grad || B == 0 || B > (1u << 20)) return VR_ERR_ARG;
    if (!c->mlpReady || c->nclusters == 0) return fail(c, VR_ERR_STATE, "vr_mlp_grad: no network or clusters");
    for (uint32_t i = 0; i < B; i++)
        if (t[i] >= 256u) return fail(c, VR_ERR_ARG, "vr_mlp_grad: target out of range");
    int rc = train_alloc(c, B);
    if (rc != VR_OK) return rc;
    HIPCHK(c, hipMemcpyAsync(c->trainX, x, (size_t)B * 36, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->trainT, t, (size_t)B * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->trainW, w, (size_t)B * 4, hipMemcpyHostToDevice, c->stream));
    const uint32_t chunks = (B + 63) / 64;
    HIPCHK(c, launch_mlp_grad(mlp_view(c), c->trainX, c->trainT, c->trainW, B, c->partial, c->chunkLoss, c->mlpG,
                              c->stream));
    HIPCHK(c, hipMemcpyAsync(grad, c->mlpG, kMlpParams * 4, hipMemcpyDeviceToHost, c->stream));
    std::vector<float> l(chunks);
    HIPCHK(c, hipMemcpyAsync(l.data(), c->chunkLoss, (size_t)chunks * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (loss) {
        double s = 0;
        for (float v : l) s += v;
        *loss = s;
    }
    return VR_OK;
}

int vr_mlp_step(vr_ctx* c, const float* x, const uint32_t* t, const float* w, uint32_t B, float lr, double* loss)
{
    if (!c || !x || !t || !w || B == 0 || B > (1u << 20) || !(lr > 0.0f)) return VR_ERR_ARG;
    if (!c->mlpReady || c->nclusters == 0) return fail(c, VR_ERR_STATE, "vr_mlp_step: no network or clusters");
    for (uint32_t i = 0; i < B; i++)
        if (t[i] >= 256u) return fail(c, VR_ERR_ARG, "vr_mlp_step: target out of range");
    int rc = train_alloc(c, B);
    if (rc != VR_OK) return rc;
    HIPCHK(c, hipMemcpyAsync(c->trainX, x, (size_t)B * 36, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->trainT, t, (size_t)B * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->trainW, w, (size_t)B * 4, hipMemcpyHostToDevice, c->stream));
    rc = train_step(c, 0, B, lr, loss);
    if (rc == VR_OK) HIPCHK(c, hipStreamSynchronize(c->stream));
    return rc;
}

int vr_mlp_train(vr_ctx* c, const vr_ctx* from, uint32_t nsamples, float lr, uint32_t steps, double* loss)
{
    if (!c || nsamples == 0 || nsamples > (1u << 20) || steps == 0 || !(lr > 0.0f)) return VR_ERR_ARG;
    const vr_ctx* src = from ? from : c;
    if (!src->framed) return fail(c, VR_ERR_STATE, "vr_mlp_train: no rendered frame to learn from");
    if (!c->mlpReady || c->nclusters == 0 || c->clusteredLights != c->nLights)
        return fail(c, VR_ERR_STATE, "vr_mlp_train: clusters do not cover the current lights");
    if (src->nLights != c->nLights)
        return fail(c, VR_ERR_ARG, "vr_mlp_train: the sample context's light list differs from the clustered one");
    int rc = train_alloc(c, nsamples);
    if (rc != VR_OK) return rc;
    if (src != c) {
        // the sample context's frame must be complete before this stream reads it
        const hipError_t e = hipStreamSynchronize(src->stream);
        if (e != hipSuccess) return hipfail(c, e, "vr_mlp_train: sample context");
    }
    // samples from the last frame: the A planes hold its final reservoirs
    FrameArgs a = frame_args(src, nullptr);
    a.cam = src->lastCam;
    a.L.n = src->nLights;
    HIPCHK(c, launch_train_gather(a, c->clusterOf, nsamples, (uint32_t)src->frameNo * 0x9e3779b9u + 0x7f4a7c15u,
                                  c->trainX, c->trainT, c->trainW, c->stream));
    // `steps` minibatches of consecutive samples (each spread over the image)
    const uint32_t nb = steps < nsamples ? steps : nsamples;
    for (uint32_t s = 0; s < nb; s++) {
        const uint32_t first = (uint32_t)((uint64_t)nsamples * s / nb);
        const uint32_t last = (uint32_t)((uint64_t)nsamples * (s + 1) / nb);
        rc = train_step(c, first, last - first, lr, (s + 1 == nb) ? loss : nullptr);
        if (rc != VR_OK) return rc;
    }
    return VR_OK;
}

int vr_read_candidates(vr_ctx* c, uint32_t* idx, float* pdf)
{
    if (!c) return VR_ERR_ARG;
    if (!c->mlpReady) return fail(c, VR_ERR_STATE, "vr_read_candidates: no learned proposal");
    const size_t n = (size_t)c->cfg.width * c->cfg.height * kCandidates;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the 8-byte slots split into the two output arrays, through a bounded
    // host buffer (a 1080p frame holds 66 M slots)
    const size_t chunk = (size_t)1 << 20;
    std::vector<uint2> slots(n < chunk ? n : chunk);
    for (size_t i0 = 0; i0 < n; i0 += chunk) {
        const size_t m = n - i0 < chunk ? n - i0 : chunk;
        HIPCHK(c, hipMemcpy(slots.data(), c->candSlot + i0, m * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < m; i++) {
            if (idx) idx[i0 + i] = slots[i].x;
            if (pdf) std::memcpy(&pdf[i0 + i], &slots[i].y, 4);
        }
    }
    return VR_OK;
}

int vr_ground_truth(vr_ctx* c, const vr_render_params* p, uint32_t spp, uint32_t frame0, double* accum)
{
    if (!c || !p || !accum) return VR_ERR_ARG;
    if (!c->committed) return fail(c, VR_ERR_STATE, "vr_ground_truth: no world committed");
    if (!finite_camera(p)) return fail(c, VR_ERR_ARG, "vr_ground_truth: non-finite camera parameters");
    const size_t n = (size_t)c->cfg.width * c->cfg.height;
    double* dacc = nullptr;
    if (dalloc(&dacc, n * 3) != hipSuccess) return fail(c, VR_ERR_NOMEM, "vr_ground_truth: allocation failed");
    int rc = VR_OK;
    hipError_t e = hipMemcpyAsync(dacc, accum, n * 24, hipMemcpyHostToDevice, c->stream);
    // chunks of at most 64 samples per launch; chunk k starts at R0 state
    // xorshift32^(64k)(R0) and frame index frame0 + 64k
    vr_render_params q = *p;
    uint32_t done = 0;
    while (e == hipSuccess && done < spp) {
        const uint32_t k = (spp - done) < 64u ? (spp - done) : 64u;
        FrameArgs a = frame_args(c, &q);
        e = launch_ground_truth(a, k, frame0 + done, dacc, c->stream);
        for (uint32_t i = 0; i < k; i++) {
            uint32_t s = q.R0;
            s ^= s << 13, s ^= s >> 17, s ^= s << 5;
            q.R0 = s;
        }
        done += k;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(accum, dacc, n * 24, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hipfail(c, e, "vr_ground_truth");
    (void)hipStreamSynchronize(c->stream);
    dfree(dacc);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// scene motion on the device
// ---------------------------------------------------------------------------
extern "C" {

int vr_set_light_motion(vr_ctx* c, const float* velocity, uint32_t n)
{
    if (!c || (n > 0 && !velocity)) return VR_ERR_ARG;
    if (n != c->nLights)
        return fail(c, VR_ERR_ARG, "vr_set_light_motion: " + std::to_string(n) + " velocities for " +
                                       std::to_string(c->nLights) + " lights");
    std::vector<float4> vel(n);
    std::vector<uint32_t> vox(n);
    for (uint32_t i = 0; i < n; i++) {
        const float vx = velocity[3 * i], vy = velocity[3 * i + 1], vz = velocity[3 * i + 2];
        if (!std::isfinite(vx) || !std::isfinite(vy) || !std::isfinite(vz) || std::fabs(vx) > 64.0f ||
            std::fabs(vy) > 64.0f || std::fabs(vz) > 64.0f)
            return fail(c, VR_ERR_ARG, "vr_set_light_motion: velocity of light " + std::to_string(i) +
                                           " is not finite or exceeds 64 voxels per frame");
        vel[i] = make_float4(vx, vy, vz, 0.0f);
        vox[i] = c->hostLights[i].voxel;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n > c->motionCap) {
        dfree(c->lightPos), dfree(c->lightVel), dfree(c->lightVoxel), dfree(c->lightTarget);
        c->lightPos = c->lightVel = nullptr, c->lightVoxel = c->lightTarget = nullptr;
        c->motionCap = 0;
        c->motionN = 0;  // no scene step may move lights through the freed buffers
        if (dalloc(&c->lightPos, n) != hipSuccess || dalloc(&c->lightVel, n) != hipSuccess ||
            dalloc(&c->lightVoxel, n) != hipSuccess || dalloc(&c->lightTarget, n) != hipSuccess)
            return fail(c, VR_ERR_NOMEM, "vr_set_light_motion: device allocation failed");
        c->motionCap = n;
    }
    if (!c->claim && dalloc(&c->claim, kClaimBuckets) != hipSuccess)
        return fail(c, VR_ERR_NOMEM, "vr_set_light_motion: device allocation failed");
    c->motionN = n;
    if (n == 0) return VR_OK;
    // positions start at the voxel centres the light records hold
    HIPCHK(c, hipMemcpyAsync(c->lightPos, c->la, (size_t)n * 16, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->lightVel, vel.data(), (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->lightVoxel, vox.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VR_OK;
}

int vr_scene_step(vr_ctx* c, uint32_t frame, uint32_t seed, uint32_t edit_permille, uint32_t walk_steps,
                  uint32_t high_water)
{
    if (!c) return VR_ERR_ARG;
    if (!c->committed) return fail(c, VR_ERR_STATE, "vr_scene_step: no world committed");
    if (edit_permille > 1000 || high_water > c->cfg.brick_capacity || walk_steps > 4096)
        return fail(c, VR_ERR_ARG, "vr_scene_step: edit_permille <= 1000, walk_steps <= 4096 and high_water <= "
                                   "brick capacity");
    HIPCHK(c, hipSetDevice(c->device));
    c->worldEpoch++;  // the next primary pass waits for these edits
    HIPCHK(c, launch_edit_bricks(c->bricks, high_water, frame, seed, edit_permille, walk_steps, c->stream));
    if (c->motionN > 0 && c->motionN == c->nLights) {
        const LightMotionView m{c->la, c->lb, c->lightPos, c->lightVel, c->lightVoxel, c->lightTarget, c->claim,
                                c->motionN};
        HIPCHK(c, launch_move_lights(m, c->grid, c->bricks, c->cfg.brick_capacity, c->stream));
    }
    return VR_OK;
}

int vr_read_lights(vr_ctx* c, vr_light* out, uint32_t n)
{
    if (!c || (n > 0 && !out)) return VR_ERR_ARG;
    if (n != c->nLights) return fail(c, VR_ERR_ARG, "vr_read_lights: n must equal the light count");
    if (n == 0) return VR_OK;
    std::vector<float4> lb(n);
    HIPCHK(c, hipMemcpyAsync(lb.data(), c->lb, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < n; i++) {
        out[i] = c->hostLights[i];
        uint32_t packed;
        std::memcpy(&packed, &lb[i].w, 4);
        const uint32_t x = packed & 1023u, z = (packed >> 10) & 1023u, y = packed >> 20;
        out[i].position = x + z * 1024u + y * 1024u * 1024u;
    }
    return VR_OK;
}

int vr_read_bricks(vr_ctx* c, const uint32_t* ids, uint32_t n, uint32_t* out)
{
    if (!c || (n > 0 && (!ids || !out))) return VR_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < n; i++) {
        if (ids[i] >= c->cfg.brick_capacity)
            return fail(c, VR_ERR_ARG, "vr_read_bricks: brick id " + std::to_string(ids[i]) + " beyond capacity");
        HIPCHK(c, hipMemcpyAsync(out + (size_t)i * VR_BRICKSIZE, c->bricks + (size_t)ids[i] * VR_BRICKSIZE,
                                 VR_BRICKSIZE * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches of views: per stage one launch per chunk of up to 8 views
// ---------------------------------------------------------------------------
struct vr_batch {
    std::vector<vr_ctx*> views;
    int device = 0;
    hipStream_t stream = nullptr;
    static constexpr int kSlots = 4;  // frames in flight, each with its own launch arguments
    struct Slot {
        std::vector<ViewChunk> chunks;       // the frame's per-view arguments
        std::vector<ProposalArgs> props;     // one per learned view
        std::vector<void*> chunkArgs;        // kernel-argument pointers of the graph nodes
        hipGraph_t graph = nullptr;          // kept for its node handles
        hipGraphExec_t exec = nullptr;
        std::vector<hipGraphNode_t> nodes;   // in launch order
        uint64_t key = 0;
        hipEvent_t begin = nullptr, done = nullptr;
        bool inFlight = false;
    };
    Slot slot[kSlots];
    std::vector<uint32_t> learned;           // indices of the views with the learned source
    hipEvent_t viewReady = nullptr;          // orders each view's stream before the batch
    static constexpr int kChainStreams = 4;  // streams of the chains when launched directly
    hipStream_t chainStream[kChainStreams] = {};
    hipEvent_t fork = nullptr, join[kChainStreams] = {};
    uint32_t chunkViews = kChunkViews;       // views per chunk; each chunk is an independent chain
    uint32_t blocks = 0;                     // workgroups per launch (0: one per tile)
    bool graph = false;
    bool anyMis = false;
    uint64_t frameNo = 0;
    bool profiling = false;
    double profSum = 0.0;
    uint32_t profFrames = 0;
    std::string err;
};

namespace {

int bfail(vr_batch* b, int code, const std::string& msg)
{
    if (b) b->err = msg;
    return code;
}

#define BHIPCHK(b, call)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return bfail(b, VR_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// The nodes of a frame, chain by chain: a chunk's four stages in order
// (its learned views' proposals before the candidate stage).  f(k, chain,
// first, params): node k, its chunk, and whether it starts the chain.
// Chains share nothing but the device, so they may run concurrently.
template <class F>
hipError_t for_each_node(vr_batch* b, vr_batch::Slot& sl, F f)
{
    const size_t nc = sl.chunks.size();
    sl.chunkArgs.resize(nc * VR_NUM_STAGES);
    size_t k = 0;
    for (size_t c = 0; c < nc; c++) {
        bool first = true;
        for (int st = 0; st < VR_NUM_STAGES; st++) {
            if (st == VR_STAGE_INITIAL)
                for (size_t li = 0; li < sl.props.size(); li++) {
                    if (b->learned[li] / b->chunkViews != c) continue;
                    hipKernelNodeParams p;
                    hipError_t e = proposal_node(&sl.props[li], &p);
                    if (e == hipSuccess) e = f(k++, c, first, p);
                    if (e != hipSuccess) return e;
                    first = false;
                }
            hipKernelNodeParams p;
            hipError_t e = stage_views_node(st, &sl.chunks[c], b->blocks, b->anyMis, &sl.chunkArgs[c * VR_NUM_STAGES + st],
                                            &p);
            if (e == hipSuccess) e = f(k++, c, first, p);
            if (e != hipSuccess) return e;
            first = false;
        }
    }
    return hipSuccess;
}

// chunk tables of every slot for b->chunkViews views per chunk (a view's
// resolution never changes)
void build_chunks(vr_batch* b)
{
    const uint32_t n = (uint32_t)b->views.size(), cv = b->chunkViews;
    const uint32_t nc = (n + cv - 1) / cv;
    for (vr_batch::Slot& sl : b->slot) {
        sl.chunks.assign(nc, ViewChunk{});
        for (uint32_t c = 0; c < nc; c++) {
            ViewChunk& ch = sl.chunks[c];
            std::memset(&ch, 0, sizeof(ch));
            ch.n = std::min<uint32_t>(cv, n - c * cv);
            for (uint32_t j = 0; j < ch.n; j++) {
                const vr_config& cfg = b->views[c * cv + j]->cfg;
                ch.start[j + 1] = ch.start[j] + ((cfg.width + 15) / 16) * ((cfg.height + 15) / 16);
            }
        }
        sl.props.assign(b->learned.size(), ProposalArgs{});
    }
}

void drop_graph(vr_batch::Slot& sl)
{
    if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
    if (sl.graph) (void)hipGraphDestroy(sl.graph);
    sl.exec = nullptr, sl.graph = nullptr;
    sl.nodes.clear();
}

}  // namespace

extern "C" {

int vr_batch_create(vr_ctx* const* views, uint32_t n, vr_batch** out)
{
    if (!views || !out || n == 0 || n > VR_MAX_VIEWS) return VR_ERR_ARG;
    *out = nullptr;
    for (uint32_t i = 0; i < n; i++) {
        if (!views[i] || views[i]->device != views[0]->device) return VR_ERR_ARG;
        for (uint32_t j = 0; j < i; j++)
            if (views[j] == views[i]) return VR_ERR_ARG;  // a view twice would swap its buffers twice
    }
    vr_batch* b = new (std::nothrow) vr_batch();
    if (!b) return VR_ERR_NOMEM;
    b->views.assign(views, views + n);
    b->device = views[0]->device;
    for (uint32_t i = 0; i < n; i++) {
        if (views[i]->cfg.flags & VR_FLAG_MIS_SPATIAL) b->anyMis = true;
        if (views[i]->cfg.source_pdf == VR_SOURCE_LEARNED) b->learned.push_back(i);
    }
    build_chunks(b);
    hipError_t e = hipSetDevice(b->device);
    for (vr_batch::Slot& sl : b->slot) {
        if (e == hipSuccess) e = hipEventCreate(&sl.begin);
        if (e == hipSuccess) e = hipEventCreate(&sl.done);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->viewReady, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int k = 0; e == hipSuccess && k < vr_batch::kChainStreams; k++) {
        e = hipStreamCreateWithFlags(&b->chainStream[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        vr_batch_destroy(b);
        return VR_ERR_HIP;
    }
    *out = b;
    return VR_OK;
}

int vr_batch_destroy(vr_batch* b)
{
    if (!b) return VR_ERR_ARG;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (int k = 0; k < vr_batch::kChainStreams; k++)
        if (b->chainStream[k]) (void)hipStreamSynchronize(b->chainStream[k]);
    for (vr_batch::Slot& sl : b->slot) {
        drop_graph(sl);
        if (sl.begin) (void)hipEventDestroy(sl.begin);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    if (b->viewReady) (void)hipEventDestroy(b->viewReady);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (int k = 0; k < vr_batch::kChainStreams; k++) {
        if (b->chainStream[k]) (void)hipStreamDestroy(b->chainStream[k]);
        if (b->join[k]) (void)hipEventDestroy(b->join[k]);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return VR_OK;
}

const char* vr_batch_last_error(const vr_batch* b) { return b ? b->err.c_str() : "null batch"; }

int vr_batch_set_blocks(vr_batch* b, uint32_t blocks)
{
    if (!b) return VR_ERR_ARG;
    b->blocks = blocks;
    return VR_OK;
}

int vr_batch_set_chunk_views(vr_batch* b, uint32_t views_per_chunk)
{
    if (!b || views_per_chunk == 0 || views_per_chunk > (uint32_t)kChunkViews) return VR_ERR_ARG;
    if (views_per_chunk == b->chunkViews) return VR_OK;
    BHIPCHK(b, hipStreamSynchronize(b->stream));  // no frame may still read the slots
    for (vr_batch::Slot& sl : b->slot) {
        drop_graph(sl);
        sl.inFlight = false;
    }
    b->chunkViews = views_per_chunk;
    build_chunks(b);
    return VR_OK;
}

int vr_batch_set_graph(vr_batch* b, int enable)
{
    if (!b) return VR_ERR_ARG;
    b->graph = enable != 0;
    return VR_OK;
}

int vr_batch_render(vr_batch* b, const vr_render_params* params)
{
    if (!b || !params) return VR_ERR_ARG;
    const uint32_t n = (uint32_t)b->views.size();
    for (uint32_t i = 0; i < n; i++) {
        vr_ctx* c = b->views[i];
        const int rc = check_frame(c, &params[i], "vr_batch_render");
        if (rc != VR_OK) return bfail(b, rc, "view " + std::to_string(i) + ": " + c->err);
    }
    BHIPCHK(b, hipSetDevice(b->device));
    vr_batch::Slot& sl = b->slot[b->frameNo % vr_batch::kSlots];
    if (sl.inFlight) {
        // the slot's launch arguments (and graph) are free once its frame ended
        BHIPCHK(b, hipEventSynchronize(sl.done));
        if (b->profiling) {
            float ms = 0.0f;
            BHIPCHK(b, hipEventElapsedTime(&ms, sl.begin, sl.done));
            b->profSum += ms, b->profFrames++;
        }
        sl.inFlight = false;
    }
    // the views' own state (hit-buffer swap, frame counters) changes only
    // once the whole frame is queued: a failed call leaves every view as it was
    std::vector<Camera> cams(n);
    for (uint32_t i = 0, li = 0; i < n; i++) {
        vr_ctx* c = b->views[i];
        c->cur = 1 - c->cur;  // arguments of the swapped hit buffers (world.cpp:2221-2223)
        FrameArgs a = frame_args(c, &params[i]);
        c->cur = 1 - c->cur;
        cams[i] = a.cam;
        a.tileOrder = VR_TILES_LINEAR;
        sl.chunks[i / b->chunkViews].v[i % b->chunkViews] = a;
        if (c->cfg.source_pdf == VR_SOURCE_LEARNED) {
            ProposalArgs& pa = sl.props[li++];
            pa.a = a, pa.m = mlp_view(c), pa.candSlot = c->candSlot;
            pa.salt = proposal_salt(params[i].R0);
        }
        // uploads queued on the view's own stream (commits, lights) first
        BHIPCHK(b, hipEventRecord(b->viewReady, c->stream));
        BHIPCHK(b, hipStreamWaitEvent(b->stream, b->viewReady, 0));
    }
    BHIPCHK(b, hipEventRecord(sl.begin, b->stream));
    if (b->graph) {
        const uint64_t key = ((uint64_t)b->blocks << 32) | ((uint64_t)b->chunkViews << 1) | (b->anyMis ? 1u : 0u);
        if (!sl.exec || sl.key != key) {
            drop_graph(sl);
            BHIPCHK(b, hipGraphCreate(&sl.graph, 0));
            const hipError_t e = for_each_node(b, sl, [&](size_t k, size_t, bool first, const hipKernelNodeParams& p) {
                hipGraphNode_t node = nullptr;
                const hipGraphNode_t* dep = first ? nullptr : &sl.nodes[k - 1];
                const hipError_t r = hipGraphAddKernelNode(&node, sl.graph, dep, first ? 0 : 1, &p);
                if (r == hipSuccess) sl.nodes.push_back(node);
                return r;
            });
            if (e != hipSuccess) {
                drop_graph(sl);
                return bfail(b, VR_ERR_HIP, std::string("vr_batch_render: graph build: ") + hipGetErrorString(e));
            }
            const hipError_t e2 = hipGraphInstantiate(&sl.exec, sl.graph, nullptr, nullptr, 0);
            if (e2 != hipSuccess) {
                sl.exec = nullptr;
                drop_graph(sl);
                return bfail(b, VR_ERR_HIP, std::string("vr_batch_render: instantiate: ") + hipGetErrorString(e2));
            }
            sl.key = key;
        } else {
            // this frame's arguments into the slot's nodes (the slot's
            // previous launch has finished)
            BHIPCHK(b, for_each_node(b, sl, [&](size_t k, size_t, bool, const hipKernelNodeParams& p) {
                        return hipGraphExecKernelNodeSetParams(sl.exec, sl.nodes[k], &p);
                    }));
        }
        BHIPCHK(b, hipGraphLaunch(sl.exec, b->stream));
    } else {
        // chains on their own streams (round robin), forked from and joined
        // back into the batch stream
        const size_t nc = sl.chunks.size();
        const int ns = (int)std::min<size_t>(nc, vr_batch::kChainStreams);
        if (ns > 1) {
            BHIPCHK(b, hipEventRecord(b->fork, b->stream));
            for (int k = 0; k < ns; k++) BHIPCHK(b, hipStreamWaitEvent(b->chainStream[k], b->fork, 0));
        }
        BHIPCHK(b, for_each_node(b, sl, [&](size_t, size_t c, bool, const hipKernelNodeParams& p) {
                    hipStream_t st = ns > 1 ? b->chainStream[c % vr_batch::kChainStreams] : b->stream;
                    return hipLaunchKernel(p.func, p.gridDim, p.blockDim, p.kernelParams, 0, st);
                }));
        if (ns > 1)
            for (int k = 0; k < ns; k++) {
                BHIPCHK(b, hipEventRecord(b->join[k], b->chainStream[k]));
                BHIPCHK(b, hipStreamWaitEvent(b->stream, b->join[k], 0));
            }
    }
    BHIPCHK(b, hipEventRecord(sl.done, b->stream));
    sl.inFlight = true;
    // later work of each view (read-backs, commits) follows the frame
    for (vr_ctx* c : b->views) BHIPCHK(b, hipStreamWaitEvent(c->stream, sl.done, 0));
    for (uint32_t i = 0; i < n; i++) {
        vr_ctx* c = b->views[i];
        c->cur = 1 - c->cur;  // primaryHitBuffer swap, world.cpp:2221-2223
        c->lastCam = cams[i];
        c->lastSlot = -1;
        c->initRecorded = false;  // this frame's passes ran on the batch's streams
        c->frameNo++;
        c->framed = true;
    }
    b->frameNo++;
    return VR_OK;
}

int vr_batch_sync(vr_batch* b)
{
    if (!b) return VR_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return VR_OK;
}

int vr_batch_profile_begin(vr_batch* b)
{
    if (!b) return VR_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    for (vr_batch::Slot& sl : b->slot) sl.inFlight = false;
    b->profSum = 0.0, b->profFrames = 0;
    b->profiling = true;
    return VR_OK;
}

int vr_batch_profile_read(vr_batch* b, double* ms_sum, uint32_t* frames)
{
    if (!b || !ms_sum || !frames) return VR_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    for (vr_batch::Slot& sl : b->slot)
        if (sl.inFlight) {
            if (b->profiling) {
                float ms = 0.0f;
                BHIPCHK(b, hipEventElapsedTime(&ms, sl.begin, sl.done));
                b->profSum += ms, b->profFrames++;
            }
            sl.inFlight = false;
        }
    *ms_sum = b->profSum;
    *frames = b->profFrames;
    return VR_OK;
}

}  // extern "C"
