// mlp.hip -- the ladder network (include/drift_mlp.h): forward, backward and Adam kernels.
//
// Sizes are tiny (39,328 parameters, batches of a few thousand), so each layer is one kernel with
// one thread per output element and a fixed-order inner loop: results do not depend on the launch
// shape, and gradients are plain sums over the batch (no atomics).  Weights are read as BF16 and
// widened to FP32 for the multiply-accumulate; the FP32 master copy is what Adam updates.
#include "../../include/drift_mlp.h"
#include "../../include/drift_pt.h"
#include "pt_engine.h"
#include "pt_rules.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int IN = DRIFT_MLP_IN, HID = DRIFT_MLP_HIDDEN, OUT = DRIFT_MLP_OUT;
constexpr int NP = DRIFT_MLP_PARAMS;
// parameter offsets: W1 b1 W2 b2 W3 b3 W4 b4
constexpr int OW1 = 0, OB1 = OW1 + IN * HID, OW2 = OB1 + HID, OB2 = OW2 + HID * HID;
constexpr int OW3 = OB2 + HID, OB3 = OW3 + HID * HID, OW4 = OB3 + HID, OB4 = OW4 + HID * OUT;
static_assert(OB4 + OUT == NP, "parameter count");

__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);   // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);                                    // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ inline float bf16_to_f32(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ inline float sigmoidf_(float x) { return 1.0f / (1.0f + expf(-x)); }

// pre[b][j] = bias[j] + sum_i in[b][i] W[i][j]; act = SiLU(pre) when silu
__global__ void k_linear(const float* __restrict__ in, const uint16_t* __restrict__ W,
                         const float* __restrict__ bias, int B, int din, int dout,
                         float* __restrict__ pre, float* __restrict__ act, int silu) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * dout) return;
    const int b = idx / dout, j = idx - b * dout;
    float s = bias[j];
    const float* x = in + (size_t)b * din;
    for (int i = 0; i < din; ++i) s += x[i] * bf16_to_f32(W[(size_t)i * dout + j]);
    pre[idx] = s;
    if (act) act[idx] = silu ? s * sigmoidf_(s) : s;
}

// Head: x_j = S_j / S_31 with S_j = sum_{k=1..j} (softplus(z_k) + 1e-4); with a target, also the loss
// contribution and dL/dz (the mean runs over samples and interior rungs 1..30).
__global__ void k_head(const float* __restrict__ z, const float* __restrict__ target, int B,
                       float* __restrict__ xo, float* __restrict__ gz, float* __restrict__ lossb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float* zb = z + (size_t)b * OUT;
    float d[OUT], S[OUT];
    S[0] = 0.0f;
    for (int k = 0; k < OUT; ++k) {
        const float v = zb[k];
        d[k] = (v > 20.0f ? v : log1pf(expf(v))) + 1e-4f;   // softplus, floored: strictly increasing
        if (k > 0) S[k] = S[k - 1] + d[k];
    }
    const float C = S[OUT - 1];
    for (int j = 0; j < OUT; ++j) xo[(size_t)b * OUT + j] = S[j] / C;
    if (!target) return;
    const float* tb = target + (size_t)b * OUT;
    const float norm = 1.0f / ((float)B * (float)(OUT - 2));
    float g[OUT], loss = 0.0f, gS = 0.0f;
    for (int j = 0; j < OUT; ++j) {
        const float x = S[j] / C;
        const bool interior = j > 0 && j < OUT - 1;
        const float e = interior ? x - tb[j] : 0.0f;
        loss += e * e;
        g[j] = 2.0f * e * norm;
        gS += g[j] * S[j];
    }
    lossb[b] = loss * norm;
    // dL/dd_k = (1/C) sum_{j>=k} g_j - (1/C^2) sum_j g_j S_j   (k >= 1); d_0 does not enter
    float tail = 0.0f;
    for (int k = OUT - 1; k >= 1; --k) {
        tail += g[k];
        const float gd = tail / C - gS / (C * C);
        gz[(size_t)b * OUT + k] = gd * sigmoidf_(zb[k]);
    }
    gz[(size_t)b * OUT] = 0.0f;
}

// gin[b][i] = sum_j gpre[b][j] W[i][j], then through SiLU' of the layer below (pre_below)
__global__ void k_back_input(const float* __restrict__ gpre, const uint16_t* __restrict__ W,
                             const float* __restrict__ pre_below, int B, int din, int dout,
                             float* __restrict__ gpre_below) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * din) return;
    const int b = idx / din, i = idx - b * din;
    float s = 0.0f;
    const float* g = gpre + (size_t)b * dout;
    const uint16_t* w = W + (size_t)i * dout;
    for (int j = 0; j < dout; ++j) s += g[j] * bf16_to_f32(w[j]);
    const float p = pre_below[idx];
    const float sg = sigmoidf_(p);
    gpre_below[idx] = s * sg * (1.0f + p * (1.0f - sg));
}

// gW[i][j] = sum_b in[b][i] gpre[b][j]; gb[j] = sum_b gpre[b][j]  (thread (i, j); i == din is the bias)
__global__ void k_back_weight(const float* __restrict__ in, const float* __restrict__ gpre, int B,
                              int din, int dout, float* __restrict__ gW, float* __restrict__ gb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (din + 1) * dout) return;
    const int i = idx / dout, j = idx - i * dout;
    float s = 0.0f;
    if (i < din) {
        for (int b = 0; b < B; ++b) s += in[(size_t)b * din + i] * gpre[(size_t)b * dout + j];
        gW[(size_t)i * dout + j] = s;
    } else {
        for (int b = 0; b < B; ++b) s += gpre[(size_t)b * dout + j];
        gb[j] = s;
    }
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, uint16_t* __restrict__ wbf, int n, float lr,
                       float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    w[i] = wi;
    wbf[i] = f32_to_bf16(wi);
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wbf, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wbf[i] = f32_to_bf16(w[i]);
}

inline int blocks(int n, int t = 256) { return (n + t - 1) / t; }

}  // namespace

struct drift_mlp {
    float* w = nullptr;       // FP32 master
    uint16_t* wbf = nullptr;  // BF16 copy used by the forward pass
    float* m = nullptr;
    float* v = nullptr;
    float* g = nullptr;
    float* work = nullptr;    // activations and gradients for up to `cap` samples
    int cap = 0;
    int64_t step = 0;
    hipStream_t s = nullptr;
};

namespace {

#define MLP_HIP(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return pt_fail(e_ == hipErrorOutOfMemory ? DRIFT_PT_ENOMEM : DRIFT_PT_EHIP,       \
                           (std::string(#call) + ": " + hipGetErrorString(e_)).c_str());      \
    } while (0)

// work layout per batch of B: X(B*IN) p1 a1 p2 a2 p3 a3 (B*HID each) z(B*OUT) x(B*OUT) T(B*OUT)
// gz(B*OUT) g3 g2 g1 (B*HID each) loss(B)
size_t work_floats(int B) { return (size_t)B * (IN + 6 * HID + 4 * OUT + 3 * HID + 1); }

int ensure_cap(drift_mlp* net, int B) {
    if (B <= net->cap) return 0;
    if (net->work) (void)hipFree(net->work);
    net->work = nullptr;
    MLP_HIP(hipMalloc(&net->work, work_floats(B) * sizeof(float)));
    net->cap = B;
    return 0;
}

struct Views {
    float *X, *p1, *a1, *p2, *a2, *p3, *a3, *z, *x, *T, *gz, *g3, *g2, *g1, *loss;
};

Views views(drift_mlp* net, int B) {
    Views v;
    float* p = net->work;
    auto take = [&](size_t n) { float* r = p; p += n; return r; };
    v.X = take((size_t)B * IN);
    v.p1 = take((size_t)B * HID); v.a1 = take((size_t)B * HID);
    v.p2 = take((size_t)B * HID); v.a2 = take((size_t)B * HID);
    v.p3 = take((size_t)B * HID); v.a3 = take((size_t)B * HID);
    v.z = take((size_t)B * OUT); v.x = take((size_t)B * OUT); v.T = take((size_t)B * OUT);
    v.gz = take((size_t)B * OUT);
    v.g3 = take((size_t)B * HID); v.g2 = take((size_t)B * HID); v.g1 = take((size_t)B * HID);
    v.loss = take((size_t)B);
    return v;
}

void forward(drift_mlp* net, const Views& v, int B) {
    hipStream_t s = net->s;
    k_linear<<<blocks(B * HID), 256, 0, s>>>(v.X, net->wbf + OW1, net->w + OB1, B, IN, HID, v.p1, v.a1, 1);
    k_linear<<<blocks(B * HID), 256, 0, s>>>(v.a1, net->wbf + OW2, net->w + OB2, B, HID, HID, v.p2, v.a2, 1);
    k_linear<<<blocks(B * HID), 256, 0, s>>>(v.a2, net->wbf + OW3, net->w + OB3, B, HID, HID, v.p3, v.a3, 1);
    k_linear<<<blocks(B * OUT), 256, 0, s>>>(v.a3, net->wbf + OW4, net->w + OB4, B, HID, OUT, v.z, nullptr, 0);
}

void backward(drift_mlp* net, const Views& v, int B) {
    hipStream_t s = net->s;
    float* g = net->g;
    k_back_weight<<<blocks((HID + 1) * OUT), 256, 0, s>>>(v.a3, v.gz, B, HID, OUT, g + OW4, g + OB4);
    k_back_input<<<blocks(B * HID), 256, 0, s>>>(v.gz, net->wbf + OW4, v.p3, B, HID, OUT, v.g3);
    k_back_weight<<<blocks((HID + 1) * HID), 256, 0, s>>>(v.a2, v.g3, B, HID, HID, g + OW3, g + OB3);
    k_back_input<<<blocks(B * HID), 256, 0, s>>>(v.g3, net->wbf + OW3, v.p2, B, HID, HID, v.g2);
    k_back_weight<<<blocks((HID + 1) * HID), 256, 0, s>>>(v.a1, v.g2, B, HID, HID, g + OW2, g + OB2);
    k_back_input<<<blocks(B * HID), 256, 0, s>>>(v.g2, net->wbf + OW2, v.p1, B, HID, HID, v.g1);
    k_back_weight<<<blocks((IN + 1) * HID), 256, 0, s>>>(v.X, v.g1, B, IN, HID, g + OW1, g + OB1);
}

int loss_and_grads(drift_mlp* net, const float* X, const float* T, int B, float* loss) {
    if (!net || !X || !T || B < 1) return pt_fail(DRIFT_PT_EINVAL, "bad arguments to the ladder network");
    int rc = ensure_cap(net, B);
    if (rc) return rc;
    Views v = views(net, B);
    MLP_HIP(hipMemcpyAsync(v.X, X, sizeof(float) * (size_t)B * IN, hipMemcpyHostToDevice, net->s));
    MLP_HIP(hipMemcpyAsync(v.T, T, sizeof(float) * (size_t)B * OUT, hipMemcpyHostToDevice, net->s));
    forward(net, v, B);
    k_head<<<blocks(B, 64), 64, 0, net->s>>>(v.z, v.T, B, v.x, v.gz, v.loss);
    backward(net, v, B);
    MLP_HIP(hipGetLastError());
    std::vector<float> lb((size_t)B);
    MLP_HIP(hipMemcpyAsync(lb.data(), v.loss, sizeof(float) * B, hipMemcpyDeviceToHost, net->s));
    MLP_HIP(hipStreamSynchronize(net->s));
    double tot = 0.0;
    for (float x : lb) tot += x;
    if (loss) *loss = (float)tot;
    return 0;
}

}  // namespace

extern "C" {

int drift_mlp_create(uint64_t seed, drift_mlp** out) {
    pt_clear_error();
    if (!out) return pt_fail(DRIFT_PT_EINVAL, "null handle pointer");
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) {
        (void)hipGetLastError();
        return pt_fail(DRIFT_PT_ENODEV, "no HIP device visible (GPU engine not built for this host)");
    }
    drift_mlp* net = new drift_mlp();
    auto cleanup = [&](int rc) { drift_mlp_destroy(net); return rc; };
    if (hipStreamCreate(&net->s) != hipSuccess ||
        hipMalloc(&net->w, sizeof(float) * NP) != hipSuccess ||
        hipMalloc(&net->wbf, sizeof(uint16_t) * NP) != hipSuccess ||
        hipMalloc(&net->m, sizeof(float) * NP) != hipSuccess ||
        hipMalloc(&net->v, sizeof(float) * NP) != hipSuccess ||
        hipMalloc(&net->g, sizeof(float) * NP) != hipSuccess)
        return cleanup(pt_fail(DRIFT_PT_ENOMEM, "ladder network allocation failed"));
    // uniform(-1, 1) * sqrt(6 / fan_in) from the Philox stream of the seed; zero biases
    std::vector<float> w((size_t)NP, 0.0f);
    const int layers[4][3] = {{OW1, IN, HID}, {OW2, HID, HID}, {OW3, HID, HID}, {OW4, HID, OUT}};
    for (int L = 0; L < 4; ++L) {
        const float scale = std::sqrt(6.0f / (float)layers[L][1]);
        const int cnt = layers[L][1] * layers[L][2];
        for (int i = 0; i < cnt; ++i) {
            const pt_u4 r = pt_philox((uint32_t)i >> 2, (uint32_t)L, 0u, 0x4D4C50u, (uint32_t)seed,
                                      (uint32_t)(seed >> 32));
            const uint32_t u = pt_word(r, i & 3);
            w[(size_t)layers[L][0] + i] = scale * (2.0f * ((float)(u >> 8) * (1.0f / 16777216.0f)) - 1.0f);
        }
    }
    int rc = drift_mlp_set_params(net, w.data());
    if (rc) return cleanup(rc);
    if (hipMemset(net->m, 0, sizeof(float) * NP) != hipSuccess ||
        hipMemset(net->v, 0, sizeof(float) * NP) != hipSuccess)
        return cleanup(pt_fail(DRIFT_PT_EHIP, "ladder network initialisation failed"));
    *out = net;
    return 0;
}

int drift_mlp_destroy(drift_mlp* net) {
    if (!net) return 0;
    if (net->w) (void)hipFree(net->w);
    if (net->wbf) (void)hipFree(net->wbf);
    if (net->m) (void)hipFree(net->m);
    if (net->v) (void)hipFree(net->v);
    if (net->g) (void)hipFree(net->g);
    if (net->work) (void)hipFree(net->work);
    if (net->s) (void)hipStreamDestroy(net->s);
    delete net;
    return 0;
}

int drift_mlp_get_params(const drift_mlp* net, float* params) {
    if (!net || !params) return pt_fail(DRIFT_PT_EINVAL, "bad arguments to drift_mlp_get_params");
    MLP_HIP(hipMemcpy(params, net->w, sizeof(float) * NP, hipMemcpyDeviceToHost));
    return 0;
}

int drift_mlp_set_params(drift_mlp* net, const float* params) {
    if (!net || !params) return pt_fail(DRIFT_PT_EINVAL, "bad arguments to drift_mlp_set_params");
    MLP_HIP(hipMemcpy(net->w, params, sizeof(float) * NP, hipMemcpyHostToDevice));
    k_to_bf16<<<blocks(NP), 256, 0, net->s>>>(net->w, net->wbf, NP);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipStreamSynchronize(net->s));
    return 0;
}

int drift_mlp_forward(drift_mlp* net, const float* X, int32_t batch, float* x_out) {
    pt_clear_error();
    if (!net || !X || !x_out || batch < 1) return pt_fail(DRIFT_PT_EINVAL, "bad arguments to drift_mlp_forward");
    int rc = ensure_cap(net, batch);
    if (rc) return rc;
    Views v = views(net, batch);
    MLP_HIP(hipMemcpyAsync(v.X, X, sizeof(float) * (size_t)batch * IN, hipMemcpyHostToDevice, net->s));
    forward(net, v, batch);
    k_head<<<blocks(batch, 64), 64, 0, net->s>>>(v.z, nullptr, batch, v.x, nullptr, nullptr);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(x_out, v.x, sizeof(float) * (size_t)batch * OUT, hipMemcpyDeviceToHost, net->s));
    MLP_HIP(hipStreamSynchronize(net->s));
    return 0;
}

int drift_mlp_gradients(drift_mlp* net, const float* X, const float* target, int32_t batch,
                        float* grads, float* loss) {
    pt_clear_error();
    if (!grads) return pt_fail(DRIFT_PT_EINVAL, "null gradient buffer");
    int rc = loss_and_grads(net, X, target, batch, loss);
    if (rc) return rc;
    MLP_HIP(hipMemcpy(grads, net->g, sizeof(float) * NP, hipMemcpyDeviceToHost));
    return 0;
}

int drift_mlp_train_step(drift_mlp* net, const float* X, const float* target, int32_t batch,
                         float lr, float* loss) {
    pt_clear_error();
    int rc = loss_and_grads(net, X, target, batch, loss);
    if (rc) return rc;
    net->step += 1;
    const float c1 = 1.0f - std::pow(0.9f, (float)net->step);
    const float c2 = 1.0f - std::pow(0.999f, (float)net->step);
    k_adam<<<blocks(NP), 256, 0, net->s>>>(net->w, net->m, net->v, net->g, net->wbf, NP, lr, c1, c2);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipStreamSynchronize(net->s));
    return 0;
}

}  // extern "C"
