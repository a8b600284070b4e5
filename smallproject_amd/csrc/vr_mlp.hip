// vr_mlp.hip - the learned light-cluster proposal: a 9-64-64-256 MLP
// (inputs: shading position / 1024 and the 6-way one-hot voxel face;
// ReLU; softmax over light clusters) evaluated and trained with BF16 MFMA
// (v_mfma_f32_32x32x16_bf16, FP32 accumulate).
//
// Layout conventions.  Every product is written C = A * B^T with both
// operands K-contiguous ("NT"), so each lane feeds the MFMA with one 16-byte
// load per operand.  For v_mfma_f32_32x32x16_bf16 lane l supplies row l%32,
// k = 8*(l/32)..+7 of A and of B^T, and holds C(8*(q/4) + 4*(l/32) + q%4,
// l%32) in accumulator register q.
//
// Parameters: FP32 master copy (flat, same order as the oracle: W1[64][9],
// b1, W2[64][64], b2, W3[256][64], b3) plus BF16 copies for the MFMAs:
// W1b[64][16] (inputs zero-padded), W2b[64][64], W2Tb[64][64], W3b[256][64],
// W3Tb[64][256].
#include "vr_mlp.h"

namespace vr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ __bf16 to_bf16(float f)
{
    // round to nearest even: gfx950's v_cvt_pk_bf16_f32
    return static_cast<__bf16>(f);
}

__device__ __forceinline__ float from_bf16(__bf16 b)
{
    unsigned short s;
    __builtin_memcpy(&s, &b, 2);
    return __uint_as_float((uint32_t)s << 16);
}

// acc += A[32 x K] * Bt[32 x K]^T, rows of A at a (stride lda), rows of Bt
// at b (stride ldb), K a multiple of 16
__device__ __forceinline__ f32x16 mfma_nt(const __bf16* a, int lda, const __bf16* b, int ldb, int K, f32x16 acc)
{
    const int l = threadIdx.x & 63;
    const int r = l & 31, kh = (l >> 5) << 3;
    for (int k = 0; k < K; k += 16) {
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(a + r * lda + k + kh);
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(b + r * ldb + k + kh);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ int acc_row(int q) { return ((q >> 2) << 3) + (((threadIdx.x & 63) >> 5) << 2) + (q & 3); }
__device__ __forceinline__ int acc_col() { return threadIdx.x & 31; }

__device__ __forceinline__ f32x16 zero16()
{
    f32x16 z;
#pragma unroll
    for (int q = 0; q < 16; q++) z[q] = 0.0f;
    return z;
}

// LDS of one wave's 64-sample chunk (16-byte aligned rows for the operand loads)
struct __attribute__((aligned(16))) ChunkLds {
    __bf16 X[64 * 16];    // inputs, sample-major
    __bf16 H1[64 * 64];   // relu(layer 1), sample-major
    __bf16 H2[64 * 64];   // relu(layer 2), sample-major
};

// hidden layers of 64 samples: X (LDS) -> H1, H2 (LDS)
__device__ void forward_hidden(const MlpView& m, ChunkLds& s)
{
    for (int mb = 0; mb < 2; mb++)
        for (int nb = 0; nb < 2; nb++) {
            f32x16 acc = mfma_nt(s.X + mb * 32 * 16, 16, m.W1b + nb * 32 * 16, 16, 16, zero16());
            const int j = nb * 32 + acc_col();
            const float bj = m.b1[j];
#pragma unroll
            for (int q = 0; q < 16; q++) s.H1[(mb * 32 + acc_row(q)) * 64 + j] = to_bf16(fmaxf(acc[q] + bj, 0.0f));
        }
    __syncthreads();
    for (int mb = 0; mb < 2; mb++)
        for (int nb = 0; nb < 2; nb++) {
            f32x16 acc = mfma_nt(s.H1 + mb * 32 * 64, 64, m.W2b + nb * 32 * 64, 64, 64, zero16());
            const int j = nb * 32 + acc_col();
            const float bj = m.b2[j];
#pragma unroll
            for (int q = 0; q < 16; q++) s.H2[(mb * 32 + acc_row(q)) * 64 + j] = to_bf16(fmaxf(acc[q] + bj, 0.0f));
        }
    __syncthreads();
}

// logits of cluster block mb for sample block nb: lane holds sample
// nb*32 + l%32, clusters mb*32 + acc_row(q); empty clusters get -inf
// (cstart: the cluster offsets, in global memory or staged in LDS)
__device__ __forceinline__ f32x16 logits_block(const MlpView& m, const uint32_t* cstart, const ChunkLds& s, int mb,
                                               int nb)
{
    f32x16 acc = mfma_nt(m.W3b + mb * 32 * 64, 64, s.H2 + nb * 32 * 64, 64, 64, zero16());
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int c = mb * 32 + acc_row(q);
        acc[q] = (c < (int)m.nclusters && cstart[c + 1] > cstart[c]) ? acc[q] + m.b3[c] : -INFINITY;
    }
    return acc;
}

// per-sample softmax statistics (max, sum of exp) for sample block nb,
// combined over the lane pair (l, l^32) that holds one sample
__device__ __forceinline__ void softmax_stats(const MlpView& m, const uint32_t* cstart, const ChunkLds& s, int nb,
                                              float& M, float& S, float& Slo)
{
    float mx = -INFINITY, sum = 0.0f;
    for (int mb = 0; mb < 8; mb++) {
        const f32x16 z = logits_block(m, cstart, s, mb, nb);
        float bm = mx;
#pragma unroll
        for (int q = 0; q < 16; q++) bm = fmaxf(bm, z[q]);
        if (bm == -INFINITY) continue;  // nothing unmasked yet
        float bs = 0.0f;
#pragma unroll
        for (int q = 0; q < 16; q++) bs += z[q] == -INFINITY ? 0.0f : __expf(z[q] - bm);
        sum = (mx == -INFINITY ? 0.0f : sum * __expf(mx - bm)) + bs;
        mx = bm;
    }
    const float omx = __shfl_xor(mx, 32), osum = __shfl_xor(sum, 32);
    M = fmaxf(mx, omx);
    const float mine = mx == -INFINITY ? 0.0f : sum * __expf(mx - M);
    const float other = omx == -INFINITY ? 0.0f : osum * __expf(omx - M);
    S = mine + other;
    Slo = (threadIdx.x & 63) < 32 ? mine : other;  // mass held by the lower lane of the pair
}

// ---------------------------------------------------------------------------
// inference path (k_proposal, k_mlp_probs): inputs in registers, one padded
// activation buffer per wave (H1 is overwritten by H2, rows padded to 72 so the 16-byte operand
// loads of neighbouring rows fall in different LDS banks) and the output
// bias staged in LDS with -inf for clusters without lights, used as the MFMA
// accumulator's initial value
// ---------------------------------------------------------------------------
constexpr int kLdH = 72;  // padded row stride (bf16 elements, a 16-byte multiple)

struct __attribute__((aligned(16))) InferLds {
    __bf16 H[64 * kLdH];  // relu(layer 1), then relu(layer 2)
};

// A wave's 64 inputs stay in registers: lane p holds sample p's 16 padded
// inputs as two halves (k 0-7, 8-15).  The layer-1 MFMA of row block mb
// needs, in lane l, sample mb*32 + l%32's half l/32: the lane's own half or
// the other half of lane l^32.
struct InputRegs {
    bf16x8 lo, hi;
};

__device__ __forceinline__ bf16x8 shfl_xor32(bf16x8 v)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 u;
    __builtin_memcpy(&u, &v, 16);
#pragma unroll
    for (int i = 0; i < 4; i++) u[i] = (uint32_t)__shfl_xor((int)u[i], 32);
    bf16x8 o;
    __builtin_memcpy(&o, &u, 16);
    return o;
}

__device__ __forceinline__ InputRegs pack_inputs(const float (&in)[16])
{
    InputRegs r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.lo[i] = to_bf16(in[i]), r.hi[i] = to_bf16(in[8 + i]);
    return r;
}

// masked output bias: b3[c], or -inf for an empty or absent cluster
__device__ __forceinline__ void stage_bias(const MlpView& m, float* sBias)
{
    for (int c = threadIdx.x; c < kMaxClusters; c += blockDim.x)
        sBias[c] = (c < (int)m.nclusters && m.clusterStart[c + 1] > m.clusterStart[c]) ? m.b3[c] : -INFINITY;
}

__device__ void forward_infer(const MlpView& m, InferLds& s, const InputRegs& x)
{
    const int l = threadIdx.x & 63;
    const bf16x8 olo = shfl_xor32(x.lo), ohi = shfl_xor32(x.hi);
    const bf16x8 a0 = l < 32 ? x.lo : ohi;  // row block 0: samples 0-31
    const bf16x8 a1 = l < 32 ? olo : x.hi;  // row block 1: samples 32-63
    for (int mb = 0; mb < 2; mb++)
        for (int nb = 0; nb < 2; nb++) {
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(m.W1b + nb * 32 * 16 + (l & 31) * 16 + ((l >> 5) << 3));
            const f32x16 acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(mb ? a1 : a0, bv, zero16(), 0, 0, 0);
            const int j = nb * 32 + acc_col();
            const float bj = m.b1[j];
#pragma unroll
            for (int q = 0; q < 16; q++) s.H[(mb * 32 + acc_row(q)) * kLdH + j] = to_bf16(fmaxf(acc[q] + bj, 0.0f));
        }
    __syncthreads();
    // layer 2, one 32-sample row block at a time: its outputs read only its
    // own rows of layer 1, so they can replace them in place
    for (int mb = 0; mb < 2; mb++) {
        const f32x16 h0 = mfma_nt(s.H + mb * 32 * kLdH, kLdH, m.W2b, 64, 64, zero16());
        const f32x16 h1 = mfma_nt(s.H + mb * 32 * kLdH, kLdH, m.W2b + 32 * 64, 64, 64, zero16());
        __syncthreads();
        const int j = acc_col();
        const float b0 = m.b2[j], b1 = m.b2[32 + j];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            s.H[(mb * 32 + acc_row(q)) * kLdH + j] = to_bf16(fmaxf(h0[q] + b0, 0.0f));
            s.H[(mb * 32 + acc_row(q)) * kLdH + 32 + j] = to_bf16(fmaxf(h1[q] + b1, 0.0f));
        }
        __syncthreads();
    }
}

// logits of cluster block mb for sample block nb (bias as initial value)
__device__ __forceinline__ f32x16 logits_infer(const MlpView& m, const float* sBias, const InferLds& s, int mb, int nb)
{
    f32x16 acc;
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = sBias[mb * 32 + acc_row(q)];
    return mfma_nt(m.W3b + mb * 32 * 64, 64, s.H + nb * 32 * kLdH, kLdH, 64, acc);
}

// softmax statistics of sample block nb (see softmax_stats)
__device__ __forceinline__ void softmax_infer(const MlpView& m, const float* sBias, const InferLds& s, int nb, float& M,
                                              float& S, float& Slo)
{
    float mx = -INFINITY, sum = 0.0f;
#pragma unroll 1
    for (int mb = 0; mb < 8; mb++) {
        const f32x16 z = logits_infer(m, sBias, s, mb, nb);
        float bm = mx;
#pragma unroll
        for (int q = 0; q < 16; q++) bm = fmaxf(bm, z[q]);
        if (bm == -INFINITY) continue;  // nothing unmasked yet
        float bs = 0.0f;
#pragma unroll
        for (int q = 0; q < 16; q++) bs += __expf(z[q] - bm);  // exp(-inf) = 0
        sum = (mx == -INFINITY ? 0.0f : sum * __expf(mx - bm)) + bs;
        mx = bm;
    }
    const float omx = __shfl_xor(mx, 32), osum = __shfl_xor(sum, 32);
    M = fmaxf(mx, omx);
    const float mine = mx == -INFINITY ? 0.0f : sum * __expf(mx - M);
    const float other = omx == -INFINITY ? 0.0f : osum * __expf(omx - M);
    S = mine + other;
    Slo = (threadIdx.x & 63) < 32 ? mine : other;  // mass held by the lower lane of the pair
}

// ---------------------------------------------------------------------------
// k_proposal: per pixel, n = min(lights, candidates) <= 32 stratified cluster
// draws from the defensive mixture m(c | x) = (1 - eps) q(c | x) + eps P_c,
// where q is the network's softmax over the non-empty clusters and P_c the
// cluster's share of the total light power.  Candidate j's target is
// (j + u_j)/n of the mixture mass, u_j = hash(pixel stream, j): the n strata
// together cover the distribution once (the candidate pass reads exactly
// slots 0..n-1), so the RIS weight p-hat/p stays unbiased.  The lane
// pair of a pixel splits the CDF (the lower lane holds the clusters with
// (c & 4) == 0) and each lane walks its part once in cluster order, writing
// slot j as (cluster, m_c bits) in one 8-byte store; the candidate pass
// draws the light inside the cluster in proportion to its power, so the
// source pdf of light i is m_c * pmf_i / P_c (eps = 1 gives the plain power
// source exactly).
// ---------------------------------------------------------------------------
struct ProposalLds {
    InferLds w[4];
    float bias[kMaxClusters];
    float power[kMaxClusters];
};

__device__ __forceinline__ void proposal_tile(const FrameArgs& a, const MlpView& m, uint2* __restrict__ candSlot,
                                              uint32_t salt, int tx, int ty, ProposalLds& L)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    InferLds& s = L.w[wv];
    float* sBias = L.bias;
    float* sPower = L.power;
    stage_bias(m, sBias);
    for (int c = threadIdx.x; c < kMaxClusters; c += 256) sPower[c] = m.clusterPower[c];
    // the wave's 64 pixels: an 8x8 tile (same map as the frame kernels)
    auto pixel_of = [&](int p, int& x, int& y) {
        x = tx * 16 + (wv & 1) * 8 + (p & 7);
        y = ty * 16 + (wv >> 1) * 8 + (p >> 3);
    };
    InputRegs xin;
    {
        int x, y;
        pixel_of(lane, x, y);
        float in[16];
#pragma unroll
        for (int i = 0; i < 16; i++) in[i] = 0.0f;
        if (x < (int)a.W && y < (int)a.H) {
            const size_t idx = (size_t)y * a.W + x;
            const uint32_t hitw = a.cur.hitw[idx];
            if ((hitw & 0xffffffu) != 0u) {
                const f3 D = camera_ray(a.cam, x, y);
                const float dist = a.cur.dist[idx];
                in[0] = (D.x * dist + a.cam.E.x) * (1.0f / 1024.0f);
                in[1] = (D.y * dist + a.cam.E.y) * (1.0f / 1024.0f);
                in[2] = (D.z * dist + a.cam.E.z) * (1.0f / 1024.0f);
                in[3 + mlp_face(hitw)] = 1.0f;
            }
        }
        xin = pack_inputs(in);
    }
    __syncthreads();  // the staged bias and powers
    forward_infer(m, s, xin);
    const float keep = 1.0f - m.mix;
    const size_t plane = (size_t)a.W * a.H;
    // slots the candidate pass reads (the host rejects more than 32)
    const int ns = (int)(a.L.n < a.nCand ? a.L.n : a.nCand);
    const float invNs = ns > 0 ? 1.0f / (float)ns : 0.0f;
    const float fns = (float)ns;
    for (int nb = 0; nb < 2; nb++) {
        const int p = nb * 32 + (lane & 31);
        int x, y;
        pixel_of(p, x, y);
        const bool valid = x < (int)a.W && y < (int)a.H;
        const size_t idx = valid ? (size_t)y * a.W + x : 0;
        float M, S, Slo;
        softmax_infer(m, sBias, s, nb, M, S, Slo);
        const float kS = keep / S;  // keep * q_c = exp(z_c - M) * kS
        // mixture mass held by the lower lane
        const float massLo = Slo * kS + m.mix * m.powerLow;
        // target jitters from the pixel's proposal stream (the shading stream
        // stays untouched); both lanes of the pair see the same u_j
        const uint32_t ubase = wang_hash((valid ? a.cur.seed[idx] : 0u) ^ salt);
        auto target = [&](int j) {
            return ((float)j + (float)wang_hash(ubase + (uint32_t)j * 0x9e3779b9u) * 2.3283064365387e-10f) * invNs;
        };
        // slots [kBeg, kEnd) belong to this lane: the lower lane takes the
        // targets below massLo
        // target(j) = (j + u_j)/n with u_j <= 1 (as rounded), so every j
        // below floor(n massLo) - 2 lies strictly under massLo: the search
        // starts there and finds the same first target at or above it
        int kSplit = (int)fminf(fmaxf(floorf(massLo * fns) - 2.0f, 0.0f), fns);
        while (kSplit < ns && target(kSplit) < massLo) kSplit++;
        const bool upper = lane >= 32;
        const int kBeg = upper ? kSplit : 0, kEnd = upper ? ns : kSplit;
        const float base = upper ? massLo : 0.0f;
        int k = kBeg;
        float tk = k < kEnd ? target(k) : 2.0f;  // next target of this lane
        uint2* slot = candSlot + (size_t)kBeg * plane + idx;  // slot k of the pixel
        float cdf = base;
        uint32_t lastC = 0xffffffffu;
        float lastM = 0.0f;
#pragma unroll 1
        for (int mb = 0; mb < 8; mb++) {
            const f32x16 z = logits_infer(m, sBias, s, mb, nb);
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const uint32_t c = (uint32_t)(mb * 32 + acc_row(q));
                const float mc = __expf(z[q] - M) * kS + m.mix * sPower[c];
                cdf = cdf + mc;
                if (mc > 0.0f) {
                    while (tk < cdf) {
                        if (valid) *slot = make_uint2(c, __float_as_uint(mc));
                        slot += plane;
                        k++;
                        tk = k < kEnd ? target(k) : 2.0f;
                    }
                    lastC = c, lastM = mc;
                }
            }
        }
        // rounding can leave targets just above the walked mass: they take
        // the last cluster with mass; a lane without any mass marks them
        // empty (candidate skipped with weight 0)
        for (; k < kEnd; k++, slot += plane)
            if (valid) *slot = make_uint2(lastC, __float_as_uint(lastM));
    }
}

__global__ __launch_bounds__(256) void k_proposal(FrameArgs a, MlpView m, uint2* __restrict__ candSlot, uint32_t salt)
{
    __shared__ ProposalLds L;
    int tx, ty;
    tile_of_block(a.tileOrder, tx, ty);
    proposal_tile(a, m, candSlot, salt, tx, ty, L);
}


// ---------------------------------------------------------------------------
// probabilities for explicit inputs (test and evaluation hook), computed as
// k_proposal computes them
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_mlp_probs(MlpView m, const float* __restrict__ x, uint32_t B,
                                                  float* __restrict__ out)
{
    __shared__ InferLds s;
    __shared__ float sBias[kMaxClusters];
    const int lane = threadIdx.x & 63;
    const uint32_t base = blockIdx.x * 64;
    stage_bias(m, sBias);
    InputRegs xin;
    {
        const uint32_t b = base + lane;
        float in[16];
#pragma unroll
        for (int i = 0; i < 16; i++) in[i] = b < B && i < 9 ? x[(size_t)b * 9 + i] : 0.0f;
        xin = pack_inputs(in);
    }
    __syncthreads();
    forward_infer(m, s, xin);
    for (int nb = 0; nb < 2; nb++) {
        float M, S, Slo;
        softmax_infer(m, sBias, s, nb, M, S, Slo);
        const uint32_t b = base + nb * 32 + (lane & 31);
#pragma unroll 1
        for (int mb = 0; mb < 8; mb++) {
            const f32x16 z = logits_infer(m, sBias, s, mb, nb);
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int c = mb * 32 + acc_row(q);
                if (b < B) out[(size_t)b * 256 + c] = __expf(z[q] - M) / S;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// training: per 64-sample chunk, forward + softmax cross-entropy backward;
// writes the chunk's gradient (NPARAM floats) and weighted loss
// ---------------------------------------------------------------------------
struct __attribute__((aligned(16))) GradLds {
    ChunkLds f;
    __bf16 XT[32 * 64];    // inputs, feature-major (rows 16..31 zero)
    __bf16 H1T[64 * 64];   // feature-major copies for weight gradients
    __bf16 H2T[64 * 64];
    __bf16 dZT[256 * 64];  // dL/dz, cluster-major
    __bf16 dZ[64 * 256];   // dL/dz, sample-major
    __bf16 dH2[64 * 64];
    __bf16 dH2T[64 * 64];
    __bf16 dH1T[64 * 64];
};

// Four waves per 64-sample chunk share its LDS: every matrix product of the
// backward pass is split into 32x32 output tiles dealt to the waves, with
// a barrier between dependent products.  The softmax statistics of a
// sample combine the two waves that hold its two halves of the clusters.
__global__ __launch_bounds__(256) void k_mlp_grad(MlpView m, const float* __restrict__ x, const uint32_t* __restrict__ tgt,
                                                  const float* __restrict__ wgt, uint32_t B, float invB,
                                                  float* __restrict__ partial, float* __restrict__ lossOut)
{
    __shared__ GradLds s;
    __shared__ float stats[4][2][64];  // per wave: max and sum of exp of its cluster half, per sample
    __shared__ float lossW[4];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const uint32_t base = blockIdx.x * 64;
    float* g = partial + (size_t)blockIdx.x * kMlpParams;
    {
        // inputs: sample tid & 63, features 4 * wv .. +3 (9 real, zero-padded to 16)
        const int smp = tid & 63;
        const uint32_t b = base + smp;
        for (int i = wv * 4; i < wv * 4 + 4; i++) {
            const __bf16 v = to_bf16(b < B && i < 9 ? x[(size_t)b * 9 + i] : 0.0f);
            s.f.X[smp * 16 + i] = v;
            s.XT[i * 64 + smp] = v;
            s.XT[(16 + i) * 64 + smp] = to_bf16(0.0f);
        }
    }
    __syncthreads();
    const int tm = wv >> 1, tn = wv & 1;  // this wave's tile of a 64x64 product
    {
        const f32x16 acc = mfma_nt(s.f.X + tm * 32 * 16, 16, m.W1b + tn * 32 * 16, 16, 16, zero16());
        const int j = tn * 32 + acc_col();
        const float bj = m.b1[j];
#pragma unroll
        for (int q = 0; q < 16; q++) s.f.H1[(tm * 32 + acc_row(q)) * 64 + j] = to_bf16(fmaxf(acc[q] + bj, 0.0f));
    }
    __syncthreads();
    {
        const f32x16 acc = mfma_nt(s.f.H1 + tm * 32 * 64, 64, m.W2b + tn * 32 * 64, 64, 64, zero16());
        const int j = tn * 32 + acc_col();
        const float bj = m.b2[j];
#pragma unroll
        for (int q = 0; q < 16; q++) s.f.H2[(tm * 32 + acc_row(q)) * 64 + j] = to_bf16(fmaxf(acc[q] + bj, 0.0f));
    }
    __syncthreads();
    // feature-major copies (wave wv: features 16 wv .. +15)
    for (int i = wv * 16; i < wv * 16 + 16; i++) {
        s.H1T[i * 64 + lane] = s.f.H1[lane * 64 + i];
        s.H2T[i * 64 + lane] = s.f.H2[lane * 64 + i];
    }
    // logits: wave wv holds sample block nb = wv >> 1, cluster blocks
    // 4 (wv & 1) .. +3; statistics over its half, then over both halves
    const int nb = wv >> 1, mb0 = (wv & 1) * 4;
    f32x16 z[4];
    float mx = -INFINITY;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        z[k] = logits_block(m, m.clusterStart, s.f, mb0 + k, nb);
#pragma unroll
        for (int q = 0; q < 16; q++) mx = fmaxf(mx, z[k][q]);
    }
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    float sum = 0.0f;
    if (mx != -INFINITY) {
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int q = 0; q < 16; q++) sum += z[k][q] == -INFINITY ? 0.0f : __expf(z[k][q] - mx);
    }
    sum += __shfl_xor(sum, 32);
    if (lane < 32) stats[wv][0][lane] = mx, stats[wv][1][lane] = sum;
    __syncthreads();
    {
        const int other = wv ^ 1, col = lane & 31;
        const float omx = stats[other][0][col], osum = stats[other][1][col];
        const float M = fmaxf(mx, omx);
        const float S = (mx == -INFINITY ? 0.0f : sum * __expf(mx - M)) + (omx == -INFINITY ? 0.0f : osum * __expf(omx - M));
        const int smp = nb * 32 + col;
        const uint32_t b = base + smp;
        const float w = b < B ? wgt[b] * invB : 0.0f;
        const uint32_t t = b < B ? tgt[b] : 0u;
        float loss = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int c = (mb0 + k) * 32 + acc_row(q);
                const float p = z[k][q] == -INFINITY ? 0.0f : __expf(z[k][q] - M) / S;
                const float d = w * (p - (c == (int)t ? 1.0f : 0.0f));
                const __bf16 db = to_bf16(d);
                s.dZT[c * 64 + smp] = db;
                s.dZ[smp * 256 + c] = db;
                if (c == (int)t && w > 0.0f) loss += -w * (z[k][q] - M - __logf(S));
            }
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
        if (lane == 0) lossW[wv] = loss;
    }
    __syncthreads();
    // dW3 = dZ^T H2 (16 tiles, 4 per wave) and db3 (one cluster per thread)
    for (int k = wv * 4; k < wv * 4 + 4; k++) {
        const int mb = k >> 1, nbw = k & 1;
        const f32x16 acc = mfma_nt(s.dZT + mb * 32 * 64, 64, s.H2T + nbw * 32 * 64, 64, 64, zero16());
#pragma unroll
        for (int q = 0; q < 16; q++) g[kOffW3 + (mb * 32 + acc_row(q)) * 64 + nbw * 32 + acc_col()] = acc[q];
    }
    {
        float sb = 0.0f;
        for (int i = 0; i < 64; i++) sb += from_bf16(s.dZT[tid * 64 + i]);
        g[kOffB3 + tid] = sb;
    }
    // dH2 = dZ W3, masked by relu'(layer 2)
    {
        const f32x16 acc = mfma_nt(s.dZ + tm * 32 * 256, 256, m.W3Tb + tn * 32 * 256, 256, 256, zero16());
        const int h = tn * 32 + acc_col();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int smp = tm * 32 + acc_row(q);
            const float d = from_bf16(s.f.H2[smp * 64 + h]) > 0.0f ? acc[q] : 0.0f;
            const __bf16 db = to_bf16(d);
            s.dH2[smp * 64 + h] = db;
            s.dH2T[h * 64 + smp] = db;
        }
    }
    __syncthreads();
    // dW2 = dH2^T H1, db2, and dH1 = dH2 W2 masked by relu'(layer 1)
    {
        const f32x16 acc = mfma_nt(s.dH2T + tm * 32 * 64, 64, s.H1T + tn * 32 * 64, 64, 64, zero16());
#pragma unroll
        for (int q = 0; q < 16; q++) g[kOffW2 + (tm * 32 + acc_row(q)) * 64 + tn * 32 + acc_col()] = acc[q];
    }
    if (wv == 0) {
        float sb = 0.0f;
        for (int i = 0; i < 64; i++) sb += from_bf16(s.dH2T[lane * 64 + i]);
        g[kOffB2 + lane] = sb;
    }
    {
        const f32x16 acc = mfma_nt(s.dH2 + tm * 32 * 64, 64, m.W2Tb + tn * 32 * 64, 64, 64, zero16());
        const int h = tn * 32 + acc_col();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int smp = tm * 32 + acc_row(q);
            const float d = from_bf16(s.f.H1[smp * 64 + h]) > 0.0f ? acc[q] : 0.0f;
            s.dH1T[h * 64 + smp] = to_bf16(d);
        }
    }
    __syncthreads();
    // dW1 = dH1^T X (9 real input columns, waves 0-1) and db1 (wave 2)
    if (wv < 2) {
        const f32x16 acc = mfma_nt(s.dH1T + wv * 32 * 64, 64, s.XT, 64, 64, zero16());
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int i = acc_col();
            if (i < 9) g[kOffW1 + (wv * 32 + acc_row(q)) * 9 + i] = acc[q];
        }
    } else if (wv == 2) {
        float sb = 0.0f;
        for (int i = 0; i < 64; i++) sb += from_bf16(s.dH1T[lane * 64 + i]);
        g[kOffB1 + lane] = sb;
    } else if (lane == 0) {
        lossOut[blockIdx.x] = lossW[0] + lossW[1] + lossW[2] + lossW[3];
    }
}

// sum the chunk gradients: grad[p] = sum_c partial[c][p].  A block takes
// 64 parameters; its four waves sum interleaved quarters of the chunks
// (chunk c to wave c % 4) and the quarters are added in wave order.
__global__ __launch_bounds__(256) void k_grad_reduce(const float* __restrict__ partial, uint32_t chunks,
                                                     float* __restrict__ grad)
{
    __shared__ float part[4][64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t p = blockIdx.x * 64u + lane;
    float sum = 0.0f;
    if (p < (uint32_t)kMlpParams)
        for (uint32_t c = wv; c < chunks; c += 4u) sum += partial[(size_t)c * kMlpParams + p];
    part[wv][lane] = sum;
    __syncthreads();
    if (wv == 0 && p < (uint32_t)kMlpParams) grad[p] = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}

// One training step's update, fused: for 64 parameters per block, the
// chunk gradients summed exactly as k_grad_reduce sums them, the Adam
// update (Kingma & Ba 2015, Algorithm 1) on the FP32 master copy, and the
// BF16 copies of the updated weights that k_mlp_cast writes (W1b's zero
// padding never changes and is not written).
__global__ __launch_bounds__(256) void k_grad_adam(const float* __restrict__ partial, uint32_t chunks,
                                                   float* __restrict__ grad, float* __restrict__ P,
                                                   float* __restrict__ mo, float* __restrict__ ve, MlpBuffers wb,
                                                   float lr, float b1, float b2, float eps, float c1, float c2)
{
    __shared__ float part[4][64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t p = blockIdx.x * 64u + lane;
    float sum = 0.0f;
    if (p < (uint32_t)kMlpParams)
        for (uint32_t c = wv; c < chunks; c += 4u) sum += partial[(size_t)c * kMlpParams + p];
    part[wv][lane] = sum;
    __syncthreads();
    if (wv != 0 || p >= (uint32_t)kMlpParams) return;
    const float gi = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
    grad[p] = gi;
    const float mi = b1 * mo[p] + (1.0f - b1) * gi;
    const float vi = b2 * ve[p] + (1.0f - b2) * gi * gi;
    mo[p] = mi, ve[p] = vi;
    const float np = P[p] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    P[p] = np;
    const __bf16 nb = to_bf16(np);
    if (p < (uint32_t)kOffB1) {
        wb.W1b[(p / 9u) * 16u + p % 9u] = nb;
    } else if (p >= (uint32_t)kOffW2 && p < (uint32_t)kOffB2) {
        const uint32_t i = p - kOffW2, o = i / 64u, k = i % 64u;
        wb.W2b[i] = nb;
        wb.W2Tb[k * 64u + o] = nb;
    } else if (p >= (uint32_t)kOffW3 && p < (uint32_t)kOffB3) {
        const uint32_t i = p - kOffW3, o = i / 64u, k = i % 64u;
        wb.W3b[i] = nb;
        wb.W3Tb[k * 256u + o] = nb;
    }
}

// BF16 copies of the master weights
__global__ __launch_bounds__(256) void k_mlp_cast(const float* __restrict__ P, __bf16* W1b, __bf16* W2b, __bf16* W2Tb,
                                                  __bf16* W3b, __bf16* W3Tb)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < 64 * 16) {
        const uint32_t o = i / 16, k = i % 16;
        W1b[i] = to_bf16(k < 9 ? P[kOffW1 + o * 9 + k] : 0.0f);
    }
    if (i < 64 * 64) {
        const uint32_t o = i / 64, k = i % 64;
        W2b[i] = to_bf16(P[kOffW2 + i]);
        W2Tb[k * 64 + o] = to_bf16(P[kOffW2 + i]);
    }
    if (i < 256 * 64) {
        const uint32_t o = i / 64, k = i % 64;
        W3b[i] = to_bf16(P[kOffW3 + i]);
        W3Tb[k * 256 + o] = to_bf16(P[kOffW3 + i]);
    }
}

// training samples from the last frame: sample k reads the pixel
// floor(hash(k ^ salt) * npx / 2^32), so any run of consecutive samples is
// spread over the whole image; target = cluster of the final reservoir's
// light, weight 1 where the reservoir holds a visible sample (W > 0)
__global__ __launch_bounds__(256) void k_train_gather(FrameArgs a, const uint32_t* __restrict__ clusterOf, uint32_t n,
                                                      uint32_t salt, float* __restrict__ x,
                                                      uint32_t* __restrict__ tgt, float* __restrict__ wgt)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t npx = a.W * a.H;
    const uint32_t idx = (uint32_t)(((uint64_t)wang_hash(k ^ salt) * npx) >> 32);
    const int px = (int)(idx % a.W), py = (int)(idx / a.W);
    const uint32_t hitw = a.cur.hitw[idx];
    float in[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t t = 0;
    float w = 0.0f;
    if ((hitw & 0xffffffu) != 0u) {
        const f3 D = camera_ray(a.cam, px, py);
        const float dist = a.cur.dist[idx];
        in[0] = (D.x * dist + a.cam.E.x) * (1.0f / 1024.0f);
        in[1] = (D.y * dist + a.cam.E.y) * (1.0f / 1024.0f);
        in[2] = (D.z * dist + a.cam.E.z) * (1.0f / 1024.0f);
        in[3 + mlp_face(hitw)] = 1.0f;
        const float W = __uint_as_float(a.A.r0[idx].z);
        const uint32_t light = a.A.r1[idx].w;
        if (W > 0.0f && light < a.L.n) {
            t = clusterOf[light];
            w = 1.0f;
        }
    }
    for (int i = 0; i < 9; i++) x[(size_t)k * 9 + i] = in[i];
    tgt[k] = t;
    wgt[k] = w;
}

// ---------------------------------------------------------------------------
hipError_t launch_proposal(const FrameArgs& a, const MlpView& m, uint2* candSlot, uint32_t salt, hipStream_t s)
{
    hipLaunchKernelGGL(k_proposal, dim3((a.W + 15) / 16, (a.H + 15) / 16), dim3(256), 0, s, a, m, candSlot, salt);
    return hipGetLastError();
}

hipError_t proposal_node(ProposalArgs* pa, hipKernelNodeParams* p)
{
    pa->ptrs[0] = &pa->a, pa->ptrs[1] = &pa->m, pa->ptrs[2] = &pa->candSlot, pa->ptrs[3] = &pa->salt;
    *p = hipKernelNodeParams{};
    p->func = reinterpret_cast<void*>(&k_proposal);
    p->gridDim = dim3((pa->a.W + 15) / 16, (pa->a.H + 15) / 16);
    p->blockDim = dim3(256);
    p->sharedMemBytes = 0;
    p->kernelParams = pa->ptrs;
    p->extra = nullptr;
    return hipSuccess;
}

hipError_t launch_mlp_probs(const MlpView& m, const float* x, uint32_t B, float* out, hipStream_t s)
{
    if (B == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_probs, dim3((B + 63) / 64), dim3(64), 0, s, m, x, B, out);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const MlpView& m, const float* x, const uint32_t* t, const float* w, uint32_t B,
                           float* partial, float* loss, float* grad, hipStream_t s)
{
    if (B == 0) return hipSuccess;
    const uint32_t chunks = (B + 63) / 64;
    hipLaunchKernelGGL(k_mlp_grad, dim3(chunks), dim3(256), 0, s, m, x, t, w, B, 1.0f / (float)B, partial, loss);
    hipLaunchKernelGGL(k_grad_reduce, dim3((kMlpParams + 63) / 64), dim3(256), 0, s, partial, chunks, grad);
    return hipGetLastError();
}

hipError_t launch_train_step(const MlpView& m, const float* x, const uint32_t* t, const float* w, uint32_t B,
                             float* partial, float* loss, float* grad, float* P, float* mo, float* ve,
                             const MlpBuffers& wb, float lr, uint32_t step, hipStream_t s)
{
    if (B == 0) return hipSuccess;
    const uint32_t chunks = (B + 63) / 64;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_mlp_grad, dim3(chunks), dim3(256), 0, s, m, x, t, w, B, 1.0f / (float)B, partial, loss);
    hipLaunchKernelGGL(k_grad_adam, dim3((kMlpParams + 63) / 64), dim3(256), 0, s, partial, chunks, grad, P, mo, ve,
                       wb, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_cast(const float* P, const MlpBuffers& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_cast, dim3(64), dim3(256), 0, s, P, b.W1b, b.W2b, b.W2Tb, b.W3b, b.W3Tb);
    return hipGetLastError();
}

hipError_t launch_train_gather(const FrameArgs& a, const uint32_t* clusterOf, uint32_t n, uint32_t salt, float* x,
                               uint32_t* t, float* w, hipStream_t s)
{
    hipLaunchKernelGGL(k_train_gather, dim3((n + 255) / 256), dim3(256), 0, s, a, clusterOf, n, salt, x, t, w);
    return hipGetLastError();
}

}  // namespace vr
