// lsgpu_mlp.hip -- the learned-gather surrogate: a small MLP trained and run on gfx950
// matrix cores (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation).
//
// Network: 32 -> 64 -> 64 -> 64 -> 64 -> 4 (ReLU between layers), the last layer
// padded to 16 outputs so every GEMM dimension is a multiple of 16.  Every product --
// forward Z = A W, weight gradient dW = A^T dZ, input gradient dA = dZ W^T -- runs
// through one hand-written MFMA GEMM whose operand addressing is given by strides,
// so transposes cost nothing.  Elementwise work (bias, ReLU, BF16 rounding, loss,
// Adam) is in small kernels.  Activations are stored in BF16 exactly where the
// products consume them; the CPU oracle rounds at the same points.
#include "lsgpu_internal.h"
#include <vector>
#include <cstring>
#include <cmath>

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace lsg {

constexpr int kLayers = 5;
constexpr int kFeat = 32;                                  // input features per row
static const int kIn[kLayers] = {kFeat, 64, 64, 64, 64};
static const int kOut[kLayers] = {64, 64, 64, 64, 16};   // last layer: 4 used + 12 zero padding
constexpr int kOutUsed = 4;

__host__ __device__ inline uint16_t f2bf(float f)
{
    // round to nearest even (no NaN inputs here)
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) { uint32_t u = (uint32_t)b << 16; float f; memcpy(&f, &u, 4); return f; }

// C[M x N] (fp32, row-major ldc) = sum_k A(m,k) B(k,n); A(m,k) = A[m*sAm + k*sAk],
// B(k,n) = B[k*sBk + n*sBn] (bf16).  One wave per 16x16 tile, four waves per block.
// Lane l feeds A row l%16 / B column l%16 with k = 4*(l/16)..+3 and receives
// C rows 4*(l/16)..+3 of column l%16 (the 16x16x16 MFMA lane layout).
__global__ __launch_bounds__(256) void k_gemm_bf16(const uint16_t* __restrict__ A, int sAm, int sAk,
                                                   const uint16_t* __restrict__ B, int sBk, int sBn,
                                                   float* __restrict__ C, int ldc, int M, int N, int K, int accumulate)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tilesN = N / 16;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= (M / 16) * tilesN) return;
    const int tm = tile / tilesN, tn = tile % tilesN;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    const uint16_t* Ar = A + (size_t)(tm * 16 + r) * sAm;
    const uint16_t* Bc = B + (size_t)(tn * 16 + r) * sBn;
    for (int k0 = 0; k0 < K; k0 += 16) {
        v4s a, b;
        for (int i = 0; i < 4; i++) {
            a[i] = (short)Ar[(size_t)(k0 + kq + i) * sAk];
            b[i] = (short)Bc[(size_t)(k0 + kq + i) * sBk];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    for (int i = 0; i < 4; i++) {
        float* c = C + (size_t)(tm * 16 + kq + i) * ldc + tn * 16 + r;
        *c = accumulate ? *c + acc[i] : acc[i];
    }
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// Z += bias (per column), A = bf16(relu(Z)) (or bf16(Z) for the last layer)
__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ bias, uint16_t* __restrict__ A, uint32_t rows,
                           uint32_t cols, int relu)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= rows * cols) return;
    const float z = Z[i] + bias[i % cols];
    Z[i] = z;
    if (A) A[i] = f2bf(relu ? (z > 0.f ? z : 0.f) : z);
}

// loss and output gradient.  y[n x 16] (first 4 used), t[n x 3] targets.
// L = mean_i [ sum_c (y_ic - t_ic)^2 + lambda (y_i3 - e_i)^2 ],  e_i = sqrt(mean_c (y_ic - t_ic)^2) (no gradient)
__global__ void k_loss_grad(const float* __restrict__ y, const float* __restrict__ t, uint32_t n, uint32_t nvalid, float lambda,
                            float* __restrict__ dy, uint16_t* __restrict__ dyb, float* __restrict__ loss_acc)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    float g[16];
    for (int c = 0; c < 16; c++) g[c] = 0.f;
    if (i < nvalid && i < n) {
        const float* yi = y + 16 * (size_t)i;
        const float* ti = t + 3 * (size_t)i;
        float se = 0.f;
        for (int c = 0; c < 3; c++) { const float d = yi[c] - ti[c]; se += d * d; g[c] = 2.f * d / (float)nvalid; }
        const float e = sqrtf(se / 3.f);
        const float de = yi[3] - e;
        g[3] = 2.f * lambda * de / (float)nvalid;
        l = se + lambda * de * de;
    }
    if (i < n)
        for (int c = 0; c < 16; c++) { dy[16 * (size_t)i + c] = g[c]; dyb[16 * (size_t)i + c] = f2bf(g[c]); }
    // block reduction of the loss (every thread of the block takes part)
    __shared__ float s[256];
    s[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w; w >>= 1) { if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w]; __syncthreads(); }
    if (threadIdx.x == 0) atomicAdd(loss_acc, s[0]);
}

// dZ = dA * relu'(Z) rounded to bf16
__global__ void k_relu_back(const float* __restrict__ dA, const float* __restrict__ Z, uint16_t* __restrict__ dZb,
                            float* __restrict__ dZ, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = Z[i] > 0.f ? dA[i] : 0.f;
    dZ[i] = g;
    dZb[i] = f2bf(g);
}

// db[c] = sum over rows of dZ[r][c]
__global__ void k_colsum(const float* __restrict__ dZ, uint32_t rows, uint32_t cols, float* __restrict__ db)
{
    const uint32_t c = blockIdx.x;
    float s = 0.f;
    for (uint32_t r = threadIdx.x; r < rows; r += 256) s += dZ[(size_t)r * cols + c];
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w; w >>= 1) { if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w]; __syncthreads(); }
    if (threadIdx.x == 0) db[c] = sh[0];
}

// Adam (Kingma & Ba 2015) on the fp32 master parameters; mask zeroes padded outputs
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       const uint8_t* __restrict__ mask, uint32_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || !mask[i]) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi; v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Inference in one launch: the whole network per 16-row tile, one tile per wave at a
// time, four waves per workgroup, workgroups striding over 64-row groups (the weights are
// loaded into LDS once per workgroup).  The weights (transposed, BF16) and biases sit in LDS; each
// wave's activations ping-pong between two LDS tiles.  The products are the same MFMA
// calls in the same order as k_gemm_bf16's (same lane layout, same K steps), and bias,
// ReLU and BF16 rounding are k_bias_act's, so the output equals the layered forward
// pass bit for bit.
// LDS rows are padded by 4 elements (8 bytes): 16 lanes reading 16 rows at one column
// then hit 16 different bank pairs instead of one bank (row pitches of 64/128 bytes)
constexpr int kPad = 4;
constexpr int kWElems = (kFeat + kPad) * 64 + 3 * (64 + kPad) * 64 + (64 + kPad) * 16;
constexpr int kBElems = 4 * 64 + 16;
constexpr uint32_t kInferBlocks = 2048;                   // workgroups of the grid-stride inference launch
struct MlpOffsets { uint32_t w[kLayers]; uint32_t b[kLayers]; };

__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ x, uint32_t n, const uint16_t* __restrict__ wb,
                                                   const float* __restrict__ p, MlpOffsets off, float* __restrict__ y)
{
    constexpr int kInL[kLayers] = {kFeat, 64, 64, 64, 64};
    constexpr int kOutL[kLayers] = {64, 64, 64, 64, 16};
    __shared__ __attribute__((aligned(16))) uint16_t wt[kWElems];   // layer l at tofs: W_l transposed, [out][in]
    __shared__ float bias[kBElems];
    __shared__ __attribute__((aligned(16))) uint16_t act[4][2][16 * (64 + kPad)];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    int tofs = 0, bofs = 0;
    for (int l = 0; l < kLayers; l++) {
        const int in = kInL[l], out = kOutL[l];
        for (int e = tid; e < in * out; e += 256) {
            const int k = e / out, c = e % out;
            wt[tofs + c * (in + kPad) + k] = wb[off.w[l] + e];
        }
        for (int c = tid; c < out; c += 256) bias[bofs + c] = p[off.b[l] + c];
        tofs += (in + kPad) * out; bofs += out;
    }
    __syncthreads();                                 // weights and biases in place
    const int r = lane & 15, kq = (lane >> 4) * 4;
    // each wave owns its activation tiles: within the tile loop only the wave's own LDS
    // writes must be visible to its lanes (no workgroup barrier)
    auto wave_sync = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // grid-stride over 64-row groups: the weights are loaded into LDS once per workgroup
    const uint32_t groups = (uint32_t)(((uint64_t)n + 63) / 64);
    for (uint32_t it = blockIdx.x; it < groups; it += gridDim.x) {
    const uint64_t row0 = ((uint64_t)it * 4u + (uint32_t)wave) * 16u;
    uint16_t* A = act[wave][0];
    uint16_t* B = act[wave][1];
    for (int e = lane; e < 16 * kFeat; e += 64) {
        const uint32_t rr = (uint32_t)(e / kFeat), k = (uint32_t)(e % kFeat);
        A[rr * (kFeat + kPad) + k] = f2bf(row0 + rr < n ? x[(size_t)(row0 + rr) * kFeat + k] : 0.f);
    }
    wave_sync();
    tofs = 0; bofs = 0;
    for (int l = 0; l < kLayers; l++) {
        const int in = kInL[l], out = kOutL[l];
        const uint16_t* W = wt + tofs;
        for (int tile = 0; tile < out / 16; tile++) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < in; k0 += 16) {
                // four consecutive BF16 values per operand: one 8-byte LDS read each
                // (offsets are multiples of 4 elements: pitches, k0 and kq are)
                const v4s a = *(const v4s*)(A + r * (in + kPad) + k0 + kq);
                const v4s b = *(const v4s*)(W + (tile * 16 + r) * (in + kPad) + k0 + kq);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            }
            const int col = tile * 16 + r;
            const float bb = bias[bofs + col];
            for (int i = 0; i < 4; i++) {
                const float z = acc[i] + bb;
                if (l < kLayers - 1) B[(kq + i) * (out + kPad) + col] = f2bf(z > 0.f ? z : 0.f);
                else if (col < kOutUsed && row0 + kq + i < n) y[(size_t)(row0 + kq + i) * kOutUsed + col] = z;
            }
        }
        wave_sync();
        uint16_t* t = A; A = B; B = t;
        tofs += (in + kPad) * out; bofs += out;
    }
    }
}

static inline unsigned nb(uint64_t n) { return (unsigned)((n + 255) / 256); }

struct MlpImpl {
    uint32_t max_batch = 0;
    size_t wofs[kLayers], bofs[kLayers], nparams = 0;
    float* d_p = nullptr; float* d_g = nullptr; float* d_m = nullptr; float* d_v = nullptr; uint8_t* d_mask = nullptr;
    uint16_t* d_wb = nullptr;                    // bf16 weights (same offsets as d_p; biases unused)
    uint16_t* d_a[kLayers + 1] = {};             // bf16 activations: a[0] = input, a[l+1] = relu(z[l])
    float* d_z[kLayers] = {};                    // fp32 pre-activations
    float* d_dz = nullptr; uint16_t* d_dzb = nullptr; float* d_da = nullptr;
    float* d_loss = nullptr;
    uint64_t step = 0;
    ~MlpImpl()
    {
        auto f = [](void* p) { if (p) (void)hipFree(p); };
        f(d_p); f(d_g); f(d_m); f(d_v); f(d_mask); f(d_wb); f(d_dz); f(d_dzb); f(d_da); f(d_loss);
        for (int l = 0; l <= kLayers; l++) f(d_a[l]);
        for (int l = 0; l < kLayers; l++) f(d_z[l]);
    }
};

static int gemm(const uint16_t* A, int sAm, int sAk, const uint16_t* B, int sBk, int sBn, float* C, int ldc, int M, int N, int K,
                int acc, hipStream_t s)
{
    const int tiles = (M / 16) * (N / 16);
    k_gemm_bf16<<<(tiles + 3) / 4, 256, 0, s>>>(A, sAm, sAk, B, sBk, sBn, C, ldc, M, N, K, acc);
    return LSG_OK;
}

// forward over rows [0, n) of the input already in d_a[0] (n multiple of 16)
static void forward(MlpImpl& M, uint32_t n, hipStream_t s)
{
    for (int l = 0; l < kLayers; l++) {
        const int in = kIn[l], out = kOut[l];
        // Z[n x out] = A[n x in] * W[in x out]  (W row-major in x out)
        gemm(M.d_a[l], in, 1, M.d_wb + M.wofs[l], out, 1, M.d_z[l], out, (int)n, out, in, 0, s);
        k_bias_act<<<nb((uint64_t)n * out), 256, 0, s>>>(M.d_z[l], M.d_p + M.bofs[l], M.d_a[l + 1], n, out, l < kLayers - 1);
    }
}

} // namespace lsg

struct lsg_mlp { lsg::MlpImpl impl; };

using namespace lsg;

extern "C" {

int lsg_mlp_create(uint32_t max_batch, lsg_mlp** out)
{
    if (!out) return fail(LSG_ERR_INVALID, "lsg_mlp_create: out is null");
    *out = nullptr;
    if (!max_batch || max_batch % 16 || max_batch > (1u << 22)) return fail(LSG_ERR_INVALID, "max_batch must be a positive multiple of 16, <= 4M");
    if (!device_ready()) return fail(LSG_ERR_NO_DEVICE, "lsg_mlp_create: no gfx950 device");
    lsg_mlp* m = new (std::nothrow) lsg_mlp;
    if (!m) return fail(LSG_ERR_OOM, "host allocation");
    MlpImpl& M = m->impl;
    M.max_batch = max_batch;
    size_t o = 0;
    for (int l = 0; l < kLayers; l++) { M.wofs[l] = o; o += (size_t)kIn[l] * kOut[l]; M.bofs[l] = o; o += kOut[l]; }
    M.nparams = o;
    hipError_t e;
#define AL(p, bytes) do { if ((e = hipMalloc(&(p), (bytes))) != hipSuccess) { delete m; return hip_fail(e, "hipMalloc", __LINE__); } } while (0)
    AL(M.d_p, 4 * o); AL(M.d_g, 4 * o); AL(M.d_m, 4 * o); AL(M.d_v, 4 * o); AL(M.d_mask, o); AL(M.d_wb, 2 * o);
    for (int l = 0; l <= kLayers; l++) AL(M.d_a[l], 2 * (size_t)max_batch * (l == 0 ? kIn[0] : kOut[l - 1]));
    for (int l = 0; l < kLayers; l++) AL(M.d_z[l], 4 * (size_t)max_batch * kOut[l]);
    AL(M.d_dz, 4 * (size_t)max_batch * 64); AL(M.d_dzb, 2 * (size_t)max_batch * 64); AL(M.d_da, 4 * (size_t)max_batch * 64);
    AL(M.d_loss, 16);
#undef AL
    std::vector<uint8_t> mask(o, 1);
    for (size_t r = 0; r < (size_t)kIn[kLayers - 1]; r++)
        for (int c = kOutUsed; c < kOut[kLayers - 1]; c++) mask[M.wofs[kLayers - 1] + r * kOut[kLayers - 1] + c] = 0;
    for (int c = kOutUsed; c < kOut[kLayers - 1]; c++) mask[M.bofs[kLayers - 1] + c] = 0;
    if ((e = hipMemcpy(M.d_mask, mask.data(), o, hipMemcpyHostToDevice)) != hipSuccess) { delete m; return hip_fail(e, "h2d", __LINE__); }
    (void)hipMemset(M.d_p, 0, 4 * o); (void)hipMemset(M.d_m, 0, 4 * o); (void)hipMemset(M.d_v, 0, 4 * o);
    *out = m;
    return LSG_OK;
}

void lsg_mlp_destroy(lsg_mlp* m) { delete m; }

uint32_t lsg_mlp_num_params(void)
{
    uint32_t o = 0;
    for (int l = 0; l < kLayers; l++) o += (uint32_t)(kIn[l] * kOut[l] + kOut[l]);
    return o;
}

int lsg_mlp_set_params(lsg_mlp* m, const float* params)
{
    if (!m || !params) return fail(LSG_ERR_INVALID, "lsg_mlp_set_params: null argument");
    MlpImpl& M = m->impl;
    std::vector<float> p(params, params + M.nparams);
    // padded outputs of the last layer stay exactly zero
    for (size_t r = 0; r < (size_t)kIn[kLayers - 1]; r++)
        for (int c = kOutUsed; c < kOut[kLayers - 1]; c++) p[M.wofs[kLayers - 1] + r * kOut[kLayers - 1] + c] = 0.f;
    for (int c = kOutUsed; c < kOut[kLayers - 1]; c++) p[M.bofs[kLayers - 1] + c] = 0.f;
    LSG_HIP(hipMemcpy(M.d_p, p.data(), 4 * M.nparams, hipMemcpyHostToDevice));
    k_to_bf16<<<nb(M.nparams), 256>>>(M.d_p, M.d_wb, (uint32_t)M.nparams);
    LSG_HIP(hipMemset(M.d_m, 0, 4 * M.nparams));
    LSG_HIP(hipMemset(M.d_v, 0, 4 * M.nparams));
    M.step = 0;
    LSG_HIP(hipDeviceSynchronize());
    return LSG_OK;
}

int lsg_mlp_get_params(const lsg_mlp* m, float* params)
{
    if (!m || !params) return fail(LSG_ERR_INVALID, "lsg_mlp_get_params: null argument");
    LSG_HIP(hipMemcpy(params, m->impl.d_p, 4 * m->impl.nparams, hipMemcpyDeviceToHost));
    return LSG_OK;
}

// x: device float[n x 32]; y: device float[n x 4]
int lsg_mlp_forward(lsg_mlp* m, const float* x, uint32_t n, float* y, void* stream)
{
    if (!m || (n && (!x || !y))) return fail(LSG_ERR_INVALID, "lsg_mlp_forward: null argument");
    MlpImpl& M = m->impl;
    hipStream_t s = (hipStream_t)stream;
    if (!n) return LSG_OK;
    MlpOffsets off;
    for (int l = 0; l < kLayers; l++) { off.w[l] = (uint32_t)M.wofs[l]; off.b[l] = (uint32_t)M.bofs[l]; }
    const uint64_t groups = ((uint64_t)n + 63) / 64;
    k_mlp_infer<<<(unsigned)std::min<uint64_t>(groups, kInferBlocks), 256, 0, s>>>(x, n, M.d_wb, M.d_p, off, y);
    LSG_HIP(hipGetLastError());
    return LSG_OK;
}

// one training step on n rows (n <= max_batch): forward, loss, backward, Adam.
// t: device float[n x 3] targets.  *loss (host) receives the mean loss when non-null
// (synchronises).  lr <= 0 computes gradients without updating (gradient checks).
int lsg_mlp_train_step(lsg_mlp* m, const float* x, const float* t, uint32_t n, float lambda, float lr, float* loss, void* stream)
{
    if (!m || !x || !t || !n) return fail(LSG_ERR_INVALID, "lsg_mlp_train_step: bad argument");
    MlpImpl& M = m->impl;
    if (n > M.max_batch) return fail(LSG_ERR_LIMIT, "lsg_mlp_train_step: n %u > max_batch %u", n, M.max_batch);
    hipStream_t s = (hipStream_t)stream;
    const uint32_t pad = (n + 15) / 16 * 16;
    if (pad != n) LSG_HIP(hipMemsetAsync(M.d_a[0], 0, 2 * (size_t)pad * kFeat, s));
    k_to_bf16<<<nb((uint64_t)n * kFeat), 256, 0, s>>>(x, M.d_a[0], n * kFeat);
    forward(M, pad, s);
    LSG_HIP(hipMemsetAsync(M.d_loss, 0, 4, s));
    LSG_HIP(hipMemsetAsync(M.d_g, 0, 4 * M.nparams, s));
    // output gradient dY (layer 4: 16 columns)
    k_loss_grad<<<nb(pad), 256, 0, s>>>(M.d_z[kLayers - 1], t, pad, n, lambda, M.d_dz, M.d_dzb, M.d_loss);
    for (int l = kLayers - 1; l >= 0; l--) {
        const int in = kIn[l], out = kOut[l];
        // dW[in x out] = A^T[in x n] dZ[n x out]:  A(m=i,k=r) = a[r*in + i], B(k=r,n=c) = dzb[r*out + c]
        gemm(M.d_a[l], 1, in, M.d_dzb, out, 1, M.d_g + M.wofs[l], out, in, out, (int)pad, 0, s);
        k_colsum<<<out, 256, 0, s>>>(M.d_dz, pad, out, M.d_g + M.bofs[l]);
        if (l == 0) break;
        // dA[n x in] = dZ[n x out] W^T[out x in]:  A(m=r,k=c) = dzb[r*out + c], B(k=c,n=i) = W[i*out + c]
        gemm(M.d_dzb, out, 1, M.d_wb + M.wofs[l], 1, out, M.d_da, in, (int)pad, in, out, 0, s);
        k_relu_back<<<nb((uint64_t)pad * in), 256, 0, s>>>(M.d_da, M.d_z[l - 1], M.d_dzb, M.d_dz, pad * in);
    }
    if (lr > 0) {
        M.step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - powf(b1, (float)M.step), c2 = 1.f - powf(b2, (float)M.step);
        k_adam<<<nb(M.nparams), 256, 0, s>>>(M.d_p, M.d_g, M.d_m, M.d_v, M.d_mask, (uint32_t)M.nparams, lr, b1, b2, 1e-8f, c1, c2);
        k_to_bf16<<<nb(M.nparams), 256, 0, s>>>(M.d_p, M.d_wb, (uint32_t)M.nparams);
    }
    LSG_HIP(hipGetLastError());
    if (loss) {
        float l = 0;
        LSG_HIP(hipMemcpyAsync(&l, M.d_loss, 4, hipMemcpyDeviceToHost, s));
        LSG_HIP(hipStreamSynchronize(s));
        *loss = l / (float)n;
    }
    return LSG_OK;
}

int lsg_mlp_get_grads(const lsg_mlp* m, float* grads)
{
    if (!m || !grads) return fail(LSG_ERR_INVALID, "lsg_mlp_get_grads: null argument");
    LSG_HIP(hipMemcpy(grads, m->impl.d_g, 4 * m->impl.nparams, hipMemcpyDeviceToHost));
    return LSG_OK;
}

// debugging/verification hook: C[MxN] = A[MxK] B[KxN] with bf16 row-major inputs on
// the device, through the same MFMA kernel (M, N, K multiples of 16)
int lsg_mlp_gemm_test(const uint16_t* A, const uint16_t* B, float* C, uint32_t M, uint32_t N, uint32_t K)
{
    if (!A || !B || !C || !M || !N || !K || M % 16 || N % 16 || K % 16) return fail(LSG_ERR_INVALID, "gemm_test: bad shape");
    gemm(A, (int)K, 1, B, (int)N, 1, C, (int)N, (int)M, (int)N, (int)K, 0, nullptr);
    LSG_HIP(hipGetLastError());
    LSG_HIP(hipDeviceSynchronize());
    return LSG_OK;
}

} // extern "C"
