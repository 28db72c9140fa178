// fa_mlp.hip -- the learned planner: an FP32 MLP 16 -> 64 -> 64 -> 64 that
// predicts log(device time) of the 64 candidate plans of an integer product
// (fa_tune.hip), trained on the GPU on autotuner timings with hand-written
// forward, backward and Adam kernels.
//
// Parameters (floats, row-major, in one device array):
//   W1[16][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
// Loss: mean over the valid (mask = 1) entries of (y - t)^2.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "fa_internal.h"
#include "../../include/flint_amd.h"

#define NI 16
#define NH 64
#define NO 64
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + NI * NH)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NH * NO)
#define NPARAM (OFF_B3 + NO)

// out[b][j] = act(sum_i in[b][i] W[i][j] + bias[j]); one thread per (b, j)
__global__ void __launch_bounds__(256) fa_k_dense_fwd(const float* __restrict__ in, int B, int nin,
                                                      const float* __restrict__ W,
                                                      const float* __restrict__ bias, int nout,
                                                      int relu, float* __restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= B * nout) return;
    const int b = e / nout, j = e - b * nout;
    float acc = bias[j];
    for (int i = 0; i < nin; i++) acc = __builtin_fmaf(in[b * nin + i], W[i * nout + j], acc);
    out[e] = (relu && acc < 0.f) ? 0.f : acc;
}

// dY = 2 (Y - T) M / count, count = sum of M (one workgroup reduction) and
// the loss sum(M (Y - T)^2) / count into loss[0]
__global__ void __launch_bounds__(1024) fa_k_loss_grad(const float* __restrict__ Y,
                                                       const float* __restrict__ T,
                                                       const float* __restrict__ M, int n,
                                                       float* __restrict__ dY,
                                                       float* __restrict__ loss)
{
    __shared__ float sc[1024], sl[1024];
    float c = 0.f, l = 0.f;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const float d = (Y[e] - T[e]) * M[e];
        c += M[e];
        l = __builtin_fmaf(d, d, l);
    }
    sc[threadIdx.x] = c;
    sl[threadIdx.x] = l;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sc[threadIdx.x] += sc[threadIdx.x + s];
            sl[threadIdx.x] += sl[threadIdx.x + s];
        }
        __syncthreads();
    }
    const float cnt = sc[0] > 0.f ? sc[0] : 1.f;
    for (int e = threadIdx.x; e < n; e += blockDim.x) dY[e] = 2.f * (Y[e] - T[e]) * M[e] / cnt;
    if (threadIdx.x == 0) loss[0] = sl[0] / cnt;
}

// dW[i][j] = sum_b in[b][i] dOut[b][j] (i < nin), db[j] = sum_b dOut[b][j]
// (row i = nin); one thread per (i, j)
__global__ void __launch_bounds__(256) fa_k_dense_bwd_w(const float* __restrict__ in, int B,
                                                        int nin, const float* __restrict__ dOut,
                                                        int nout, float* __restrict__ dW,
                                                        float* __restrict__ db)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (nin + 1) * nout) return;
    const int i = e / nout, j = e - i * nout;
    float acc = 0.f;
    if (i < nin) {
        for (int b = 0; b < B; b++) acc = __builtin_fmaf(in[b * nin + i], dOut[b * nout + j], acc);
        dW[i * nout + j] = acc;
    } else {
        for (int b = 0; b < B; b++) acc += dOut[b * nout + j];
        db[j] = acc;
    }
}

// dIn[b][i] = (sum_j dOut[b][j] W[i][j]) * (act[b][i] > 0)
__global__ void __launch_bounds__(256) fa_k_dense_bwd_x(const float* __restrict__ dOut, int B,
                                                        int nout, const float* __restrict__ W,
                                                        int nin, const float* __restrict__ act,
                                                        float* __restrict__ dIn)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= B * nin) return;
    const int b = e / nin, i = e - b * nin;
    float acc = 0.f;
    for (int j = 0; j < nout; j++) acc = __builtin_fmaf(dOut[b * nout + j], W[i * nout + j], acc);
    dIn[e] = act[e] > 0.f ? acc : 0.f;
}

// Adam (Kingma & Ba) with bias correction
__global__ void __launch_bounds__(256) fa_k_adam(float* __restrict__ P, const float* __restrict__ G,
                                                 float* __restrict__ Mo, float* __restrict__ Vo,
                                                 int n, float lr, float b1, float b2, float eps,
                                                 float c1, float c2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float g = G[e];
    const float m = b1 * Mo[e] + (1.f - b1) * g;
    const float v = b2 * Vo[e] + (1.f - b2) * g * g;
    Mo[e] = m;
    Vo[e] = v;
    P[e] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

struct fa_mlp_s {
    float* P;        // parameters
    float* G;        // gradients
    float* Mo;       // Adam moments
    float* Vo;
    int step;
    int cap;         // batch capacity of the scratch below
    float *X, *T, *Mk, *H1, *H2, *Y, *dY, *dH2, *dH1, *loss;
    hipStream_t s;
};

static unsigned nblk(int n) { return (unsigned)((n + 255) / 256); }

static int mlp_reserve(fa_mlp_s* m, int B)
{
    if (B <= m->cap) return FA_OK;
    float* bufs[] = {m->X, m->T, m->Mk, m->H1, m->H2, m->Y, m->dY, m->dH2, m->dH1};
    for (float* p : bufs)
        if (p) (void)hipFree(p);
    const size_t sz[] = {(size_t)B * NI, (size_t)B * NO, (size_t)B * NO, (size_t)B * NH,
                         (size_t)B * NH, (size_t)B * NO, (size_t)B * NO, (size_t)B * NH,
                         (size_t)B * NH};
    float** dst[] = {&m->X, &m->T, &m->Mk, &m->H1, &m->H2, &m->Y, &m->dY, &m->dH2, &m->dH1};
    for (int k = 0; k < 9; k++) {
        *dst[k] = nullptr;
        if (hipMalloc((void**)dst[k], sz[k] * sizeof(float)) != hipSuccess) {
            m->cap = 0;
            return fa_set_error(FA_E_NOMEM, "mlp scratch");
        }
    }
    m->cap = B;
    return FA_OK;
}

static void mlp_forward(fa_mlp_s* m, int B)
{
    hipLaunchKernelGGL(fa_k_dense_fwd, dim3(nblk(B * NH)), dim3(256), 0, m->s, m->X, B, NI,
                       m->P + OFF_W1, m->P + OFF_B1, NH, 1, m->H1);
    hipLaunchKernelGGL(fa_k_dense_fwd, dim3(nblk(B * NH)), dim3(256), 0, m->s, m->H1, B, NH,
                       m->P + OFF_W2, m->P + OFF_B2, NH, 1, m->H2);
    hipLaunchKernelGGL(fa_k_dense_fwd, dim3(nblk(B * NO)), dim3(256), 0, m->s, m->H2, B, NH,
                       m->P + OFF_W3, m->P + OFF_B3, NO, 0, m->Y);
}

static void mlp_backward(fa_mlp_s* m, int B)
{
    hipLaunchKernelGGL(fa_k_loss_grad, dim3(1), dim3(1024), 0, m->s, m->Y, m->T, m->Mk, B * NO,
                       m->dY, m->loss);
    hipLaunchKernelGGL(fa_k_dense_bwd_w, dim3(nblk((NH + 1) * NO)), dim3(256), 0, m->s, m->H2, B,
                       NH, m->dY, NO, m->G + OFF_W3, m->G + OFF_B3);
    hipLaunchKernelGGL(fa_k_dense_bwd_x, dim3(nblk(B * NH)), dim3(256), 0, m->s, m->dY, B, NO,
                       m->P + OFF_W3, NH, m->H2, m->dH2);
    hipLaunchKernelGGL(fa_k_dense_bwd_w, dim3(nblk((NH + 1) * NH)), dim3(256), 0, m->s, m->H1, B,
                       NH, m->dH2, NH, m->G + OFF_W2, m->G + OFF_B2);
    hipLaunchKernelGGL(fa_k_dense_bwd_x, dim3(nblk(B * NH)), dim3(256), 0, m->s, m->dH2, B, NH,
                       m->P + OFF_W2, NH, m->H1, m->dH1);
    hipLaunchKernelGGL(fa_k_dense_bwd_w, dim3(nblk((NI + 1) * NH)), dim3(256), 0, m->s, m->X, B,
                       NI, m->dH1, NH, m->G + OFF_W1, m->G + OFF_B1);
}

extern "C" void fa_mlp_destroy(fa_mlp_s* m)
{
    if (!m) return;
    float* bufs[] = {m->P, m->G, m->Mo, m->Vo, m->X, m->T, m->Mk, m->H1, m->H2, m->Y, m->dY,
                     m->dH2, m->dH1, m->loss};
    for (float* p : bufs)
        if (p) (void)hipFree(p);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
}

extern "C" int fa_mlp_param_count(void) { return NPARAM; }

// He-style uniform initialization from a seeded xorshift generator
extern "C" int fa_mlp_create(fa_mlp_s** out, uint64_t seed)
{
    if (!out) return fa_set_error(FA_E_ARG, "fa_mlp_create: null");
    if (!fa_state()) return FA_E_NODEV;
    fa_mlp_s* m = new fa_mlp_s();
    memset(m, 0, sizeof(*m));
    if (hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&m->P, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&m->G, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&m->Mo, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&m->Vo, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&m->loss, 64) != hipSuccess) {
        fa_mlp_destroy(m);
        return fa_set_error(FA_E_NOMEM, "fa_mlp_create");
    }
    std::vector<float> h(NPARAM, 0.f);
    uint64_t x = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto uni = [&]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return (double)(x >> 11) / 9007199254740992.0 * 2.0 - 1.0;
    };
    const int fan[3] = {NI, NH, NH}, offw[3] = {OFF_W1, OFF_W2, OFF_W3}, nout[3] = {NH, NH, NO};
    for (int l = 0; l < 3; l++) {
        const double sc = sqrt(6.0 / fan[l]);
        for (int k = 0; k < fan[l] * nout[l]; k++) h[offw[l] + k] = (float)(sc * uni());
    }
    std::vector<float> zero(NPARAM, 0.f);
    if (hipMemcpy(m->P, h.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->Mo, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->Vo, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        fa_mlp_destroy(m);
        return fa_set_error(FA_E_HIP, "fa_mlp_create: upload");
    }
    *out = m;
    return FA_OK;
}

extern "C" int fa_mlp_get_params(fa_mlp_s* m, float* host)
{
    if (!m || !host) return fa_set_error(FA_E_ARG, "fa_mlp_get_params");
    return hipMemcpy(host, m->P, NPARAM * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess
               ? FA_OK : fa_set_error(FA_E_HIP, "fa_mlp_get_params");
}

extern "C" int fa_mlp_set_params(fa_mlp_s* m, const float* host)
{
    if (!m || !host) return fa_set_error(FA_E_ARG, "fa_mlp_set_params");
    m->step = 0;
    std::vector<float> zero(NPARAM, 0.f);
    if (hipMemcpy(m->P, host, NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->Mo, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->Vo, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_mlp_set_params");
    return FA_OK;
}

// Y = net(X) for B rows (host arrays)
extern "C" int fa_mlp_forward(fa_mlp_s* m, const float* X, int B, float* Y)
{
    if (!m || !X || !Y || B <= 0) return fa_set_error(FA_E_ARG, "fa_mlp_forward");
    int rc = mlp_reserve(m, B);
    if (rc) return rc;
    if (hipMemcpyAsync(m->X, X, (size_t)B * NI * 4, hipMemcpyHostToDevice, m->s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_mlp_forward: upload");
    mlp_forward(m, B);
    if (hipMemcpyAsync(Y, m->Y, (size_t)B * NO * 4, hipMemcpyDeviceToHost, m->s) != hipSuccess ||
        hipStreamSynchronize(m->s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_mlp_forward");
    return FA_OK;
}

// loss and gradient (host copies) at the current parameters for B rows
extern "C" int fa_mlp_loss_grad(fa_mlp_s* m, const float* X, const float* T, const float* M,
                                int B, float* loss, float* grad)
{
    if (!m || !X || !T || !M || B <= 0) return fa_set_error(FA_E_ARG, "fa_mlp_loss_grad");
    int rc = mlp_reserve(m, B);
    if (rc) return rc;
    if (hipMemcpyAsync(m->X, X, (size_t)B * NI * 4, hipMemcpyHostToDevice, m->s) != hipSuccess ||
        hipMemcpyAsync(m->T, T, (size_t)B * NO * 4, hipMemcpyHostToDevice, m->s) != hipSuccess ||
        hipMemcpyAsync(m->Mk, M, (size_t)B * NO * 4, hipMemcpyHostToDevice, m->s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_mlp_loss_grad: upload");
    mlp_forward(m, B);
    mlp_backward(m, B);
    if (loss && hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_mlp_loss_grad: loss");
    if (grad && hipMemcpyAsync(grad, m->G, NPARAM * 4, hipMemcpyDeviceToHost, m->s) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_mlp_loss_grad: grad");
    return hipStreamSynchronize(m->s) == hipSuccess ? FA_OK : fa_set_error(FA_E_HIP, "sync");
}

// minibatch Adam on the GPU over a dataset of N rows (host arrays, taken in
// the given order; the caller shuffles); loss_out[e] = mean batch loss of
// epoch e
extern "C" int fa_mlp_train(fa_mlp_s* m, const float* X, const float* T, const float* M, int64_t N,
                            int batch, int epochs, float lr, float* loss_out)
{
    if (!m || !X || !T || !M || N <= 0 || batch <= 0 || epochs <= 0)
        return fa_set_error(FA_E_ARG, "fa_mlp_train");
    int rc = mlp_reserve(m, batch);
    if (rc) return rc;
    float *dX = nullptr, *dT = nullptr, *dM = nullptr, *dl = nullptr;
    const int64_t nb = (N + batch - 1) / batch;
    if (hipMalloc((void**)&dX, (size_t)N * NI * 4) != hipSuccess ||
        hipMalloc((void**)&dT, (size_t)N * NO * 4) != hipSuccess ||
        hipMalloc((void**)&dM, (size_t)N * NO * 4) != hipSuccess ||
        hipMalloc((void**)&dl, (size_t)nb * epochs * 4) != hipSuccess) {
        rc = fa_set_error(FA_E_NOMEM, "fa_mlp_train: dataset");
    }
    if (!rc && (hipMemcpyAsync(dX, X, (size_t)N * NI * 4, hipMemcpyHostToDevice, m->s) != hipSuccess ||
                hipMemcpyAsync(dT, T, (size_t)N * NO * 4, hipMemcpyHostToDevice, m->s) != hipSuccess ||
                hipMemcpyAsync(dM, M, (size_t)N * NO * 4, hipMemcpyHostToDevice, m->s) != hipSuccess))
        rc = fa_set_error(FA_E_HIP, "fa_mlp_train: upload");
    float* sX = m->X;
    float* sT = m->T;
    float* sM = m->Mk;
    float* sl = m->loss;
    for (int e = 0; e < epochs && !rc; e++) {
        for (int64_t k = 0; k < nb; k++) {
            const int64_t r0 = k * batch;
            const int B = (int)((N - r0 < batch) ? (N - r0) : batch);
            m->X = dX + r0 * NI;
            m->T = dT + r0 * NO;
            m->Mk = dM + r0 * NO;
            m->loss = dl + e * nb + k;
            mlp_forward(m, B);
            mlp_backward(m, B);
            m->step++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
            hipLaunchKernelGGL(fa_k_adam, dim3(nblk(NPARAM)), dim3(256), 0, m->s, m->P, m->G,
                               m->Mo, m->Vo, NPARAM, lr, b1, b2, 1e-8f, c1, c2);
        }
    }
    m->X = sX;
    m->T = sT;
    m->Mk = sM;
    m->loss = sl;
    if (!rc) {
        std::vector<float> h((size_t)nb * epochs);
        if (hipMemcpyAsync(h.data(), dl, h.size() * 4, hipMemcpyDeviceToHost, m->s) != hipSuccess ||
            hipStreamSynchronize(m->s) != hipSuccess)
            rc = fa_set_error(FA_E_HIP, "fa_mlp_train: %s", hipGetErrorString(hipGetLastError()));
        if (!rc && loss_out)
            for (int e = 0; e < epochs; e++) {
                double s = 0;
                for (int64_t k = 0; k < nb; k++) s += h[(size_t)e * nb + k];
                loss_out[e] = (float)(s / nb);
            }
    }
    if (dX) (void)hipFree(dX);
    if (dT) (void)hipFree(dT);
    if (dM) (void)hipFree(dM);
    if (dl) (void)hipFree(dl);
    return rc;
}

// learned plans: for each shape (an, bn) the valid candidate with the
// smallest predicted cost; out[2k], out[2k+1] = (np, bits)
extern "C" int fa_mlp_plan_mpn(fa_mlp_s* m, const int64_t* shapes, int n, int32_t* out)
{
    if (!m || !shapes || !out || n <= 0) return fa_set_error(FA_E_ARG, "fa_mlp_plan_mpn");
    std::vector<float> X((size_t)n * NI), Y((size_t)n * NO);
    for (int k = 0; k < n; k++) {
        int rc = fa_plan_features(shapes[2 * k], shapes[2 * k + 1], &X[(size_t)k * NI]);
        if (rc) return rc;
    }
    int rc = fa_mlp_forward(m, X.data(), n, Y.data());
    if (rc) return rc;
    for (int k = 0; k < n; k++) {
        int32_t cand[3 * NO];
        rc = fa_mpn_candidates(shapes[2 * k], shapes[2 * k + 1], cand);
        if (rc) return rc;
        int best = -1;
        for (int j = 0; j < NO; j++)
            if (cand[3 * j] && (best < 0 || Y[(size_t)k * NO + j] < Y[(size_t)k * NO + best])) best = j;
        for (int e = 0; e < 3; e++) out[3 * k + e] = best >= 0 ? cand[3 * best + e] : 0;
    }
    return FA_OK;
}
