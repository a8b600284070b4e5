// vr_mlp.h - learned light-cluster proposal: parameter layout, device views
// and launchers (kernels in vr_mlp.hip).
#pragma once
#include "vr_kernels.h"

namespace vr {

// flat parameter layout (FP32 master copy): W1[64][9], b1[64], W2[64][64],
// b2[64], W3[256][64], b3[256]
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + 64 * 9;
constexpr int kOffW2 = kOffB1 + 64;
constexpr int kOffB2 = kOffW2 + 64 * 64;
constexpr int kOffW3 = kOffB2 + 64;
constexpr int kOffB3 = kOffW3 + 256 * 64;
constexpr int kMlpParams = kOffB3 + 256;  // 21440
constexpr int kMaxClusters = 256;
constexpr int kCandidates = 32;           // candidate slots written by the proposal

// read-only view of the network and the light clusters
struct MlpView {
    const __bf16* W1b;   // [64][16]
    const __bf16* W2b;   // [64][64]
    const __bf16* W2Tb;  // [64][64]
    const __bf16* W3b;   // [256][64]
    const __bf16* W3Tb;  // [64][256]
    const float* b1;
    const float* b2;
    const float* b3;
    const uint32_t* clusterStart;  // [kMaxClusters + 1] offsets into the cluster-ordered light list
    const float* clusterPower;     // [kMaxClusters] P_c: the cluster's share of the total light power
    uint32_t nclusters;
    float mix;                     // eps of the defensive mixture (1-eps) q + eps P
    float powerLow;                // sum of P_c over the clusters with (c & 4) == 0
};

struct MlpBuffers {
    __bf16* W1b;
    __bf16* W2b;
    __bf16* W2Tb;
    __bf16* W3b;
    __bf16* W3Tb;
};

// face of a hit word for the network input: +x,+y,+z,-x,-y,-z = 0..5
// (the voxel normal points against D on the hit axis)
__device__ __forceinline__ int mlp_face(uint32_t hitw)
{
    const int side = (int)((hitw >> 24) & 3u);
    return side + (((hitw >> 26) & 1u) ? 3 : 0);
}

// salt of the proposal's target jitter for a frame with random seed R0
__host__ __device__ inline uint32_t proposal_salt(uint32_t R0) { return R0 * 0x9e3779b9u + 0x632be5abu; }

// arguments of one k_proposal launch as a graph kernel node (the node
// reads them through ptrs while it is added or updated)
struct ProposalArgs {
    FrameArgs a;
    MlpView m;
    uint2* candSlot;
    uint32_t salt;
    void* ptrs[4];
};
hipError_t proposal_node(ProposalArgs* pa, hipKernelNodeParams* p);
hipError_t launch_proposal(const FrameArgs& a, const MlpView& m, uint2* candSlot, uint32_t salt, hipStream_t s);
hipError_t launch_mlp_probs(const MlpView& m, const float* x, uint32_t B, float* out, hipStream_t s);
hipError_t launch_mlp_grad(const MlpView& m, const float* x, const uint32_t* t, const float* w, uint32_t B,
                           float* partial, float* loss, float* grad, hipStream_t s);
// one Adam step on B samples: gradient chunks, then their sum, the update
// and the BF16 copies in one kernel (the gradient bits of launch_mlp_grad,
// then Adam, then launch_mlp_cast's copies); step = Adam's t after this step
hipError_t launch_train_step(const MlpView& m, const float* x, const uint32_t* t, const float* w, uint32_t B,
                             float* partial, float* loss, float* grad, float* P, float* mo, float* ve,
                             const MlpBuffers& wb, float lr, uint32_t step, hipStream_t s);
hipError_t launch_mlp_cast(const float* P, const MlpBuffers& b, hipStream_t s);
hipError_t launch_train_gather(const FrameArgs& a, const uint32_t* clusterOf, uint32_t n, uint32_t salt, float* x,
                               uint32_t* t, float* w, hipStream_t s);

}  // namespace vr
