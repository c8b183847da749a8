// Host side of the per-quartet network (kernels in tcg_mlp.hip): inference,
// one forward/backward (for gradient checks) and Adam training on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/tcg.h"
#include "tcg_device.hpp"

using namespace tcg;

namespace {

template <class T>
struct Buf {
    T *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc((void **)&p, sizeof(T) * (n ? n : 1)); }
    hipError_t put(const T *h, size_t n) {
        hipError_t e = alloc(n);
        if (e == hipSuccess && n) e = hipMemcpy(p, h, sizeof(T) * n, hipMemcpyHostToDevice);
        return e;
    }
};

#define CHK(expr)                                \
    do {                                         \
        if ((expr) != hipSuccess) return TCG_E_DEVICE; \
    } while (0)

}  // namespace

extern "C" {

int tcg_mlp_param_count(void) { return mlp_param_count(); }

int tcg_mlp_forward(const float *params, const float *X, int n, float *Y) {
    if (!params || !X || !Y || n <= 0) return TCG_E_ARG;
    const int np = mlp_param_count();
    Buf<float> dP, dX, dY;
    CHK(dP.put(params, np));
    CHK(dX.put(X, (size_t)n * 6));
    CHK(dY.alloc((size_t)n * 2));
    CHK(launch_mlp_fwd(dP.p, dX.p, n, dY.p, 0));
    CHK(hipMemcpy(Y, dY.p, sizeof(float) * (size_t)n * 2, hipMemcpyDeviceToHost));
    return TCG_OK;
}

int tcg_mlp_grad(const float *params, const float *X, const float *T, int n, const float w[2], float *grad,
                 float *loss) {
    if (!params || !X || !T || !w || !grad || !loss || n <= 0) return TCG_E_ARG;
    const int np = mlp_param_count();
    Buf<float> dP, dX, dT, dG, dL;
    CHK(dP.put(params, np));
    CHK(dX.put(X, (size_t)n * 6));
    CHK(dT.put(T, (size_t)n * 2));
    CHK(dG.alloc(np));
    CHK(dL.alloc(1));
    CHK(hipMemset(dG.p, 0, sizeof(float) * np));
    CHK(hipMemset(dL.p, 0, sizeof(float)));
    CHK(launch_mlp_step(dP.p, dX.p, dT.p, n, 1.0f / n, w[0], w[1], dG.p, dL.p, 0));
    CHK(hipMemcpy(grad, dG.p, sizeof(float) * np, hipMemcpyDeviceToHost));
    CHK(hipMemcpy(loss, dL.p, sizeof(float), hipMemcpyDeviceToHost));
    return TCG_OK;
}

int tcg_mlp_train(float *params, const float *X, const float *T, int n, int epochs, int batch, float lr,
                  unsigned seed, const float w[2], float *loss_hist) {
    if (!params || !X || !T || !w || n <= 0 || epochs <= 0 || batch <= 0) return TCG_E_ARG;
    const int np = mlp_param_count();
    Buf<float> dP, dG, dM, dV, dX, dT, dL;
    CHK(dP.put(params, np));
    CHK(dG.alloc(np));
    CHK(dM.alloc(np));
    CHK(dV.alloc(np));
    CHK(dX.alloc((size_t)n * 6));
    CHK(dT.alloc((size_t)n * 2));
    CHK(dL.alloc(epochs));
    CHK(hipMemset(dG.p, 0, sizeof(float) * np));
    CHK(hipMemset(dM.p, 0, sizeof(float) * np));
    CHK(hipMemset(dV.p, 0, sizeof(float) * np));
    CHK(hipMemset(dL.p, 0, sizeof(float) * epochs));
    std::mt19937 rng(seed);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::vector<float> xs((size_t)n * 6), ts((size_t)n * 2);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    long long t = 0;
    const int nbatch = (n + batch - 1) / batch;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        for (int i = 0; i < n; ++i) {
            std::copy(X + (size_t)perm[i] * 6, X + (size_t)perm[i] * 6 + 6, xs.begin() + (size_t)i * 6);
            std::copy(T + (size_t)perm[i] * 2, T + (size_t)perm[i] * 2 + 2, ts.begin() + (size_t)i * 2);
        }
        CHK(hipMemcpy(dX.p, xs.data(), sizeof(float) * xs.size(), hipMemcpyHostToDevice));
        CHK(hipMemcpy(dT.p, ts.data(), sizeof(float) * ts.size(), hipMemcpyHostToDevice));
        for (int b = 0; b < nbatch; ++b) {
            const int off = b * batch, m = std::min(batch, n - off);
            ++t;
            CHK(launch_mlp_step(dP.p, dX.p + (size_t)off * 6, dT.p + (size_t)off * 2, m, 1.0f / m, w[0], w[1], dG.p,
                                dL.p + ep, 0));
            const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
            CHK(launch_mlp_adam(dP.p, dG.p, dM.p, dV.p, lr, b1, b2, eps, c1, c2, 0));
        }
    }
    CHK(hipMemcpy(params, dP.p, sizeof(float) * np, hipMemcpyDeviceToHost));
    if (loss_hist) {
        std::vector<float> l(epochs);
        CHK(hipMemcpy(l.data(), dL.p, sizeof(float) * epochs, hipMemcpyDeviceToHost));
        for (int ep = 0; ep < epochs; ++ep) loss_hist[ep] = l[ep] / nbatch;     // mean batch loss
    }
    return TCG_OK;
}

}  // extern "C"
