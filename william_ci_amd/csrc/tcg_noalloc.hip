// Coulomb and exchange from the resident ERI tensor with no device
// allocation (tcg_jk_stored_noalloc), for callers that cannot allocate:
// code under graph capture, serving loops, harnesses that forbid hipMalloc.
//
// The arithmetic is that of hartree_exchange_potential
// (src/Energies/hartree_exchange_potential.f90:30-202):
//   chemist   ERI(a,b,c,d) = (ab|cd):  J(a,b) += E P(c,d),  K(a,d) += -1/2 E P(b,c)
//   physicist ERI(a,b,c,d) = (ac|bd):  J(a,c) += E P(b,d),  K(a,d) += -1/2 E P(c,b)
// rows a <= n1 of J and K, the others zero.
//
// All device memory beyond the tensor is one file-scope array, `scratch`
// (kScratchDoubles doubles, 32 MiB per GPU, held until the process ends):
//   [0, nbas^2)                 P (nbas, nbas) column-major, uploaded per call
//   [nbas^2, nbas^2 + n1 nbas)  J rows (n1, nbas) column-major
//   [.., + n1 nbas)             K rows (n1, nbas)
// so a call needs nbas^2 + 2 n1 nbas <= kScratchDoubles (n1 = nfuc: nbas <= 2047
// for the SAO layouts here; n1 = nbas: nbas <= 1182).  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU race (one context,
// one host thread at a time, as for every entry point).  J and K are cleared
// by a fill kernel at the start of each call.
//
// Host code reaches the array only through the address hipGetSymbolAddress
// returns for the current device (scratch_ptr), resolved once per device by
// tcg_create on that device.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "tcg.h"
#include "tcg_device.hpp"

namespace tcg {

constexpr size_t kScratchDoubles = size_t(4) << 20;    // 32 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// Resolves the device address of `scratch` on the current device (once per
// device); called by tcg_create.
hipError_t noalloc_setup() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return hipSuccess;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = static_cast<double *>(p);
    return hipSuccess;
}

static double *scratch_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}

size_t noalloc_scratch_doubles() { return kScratchDoubles; }

// ---------------------------------------------------------------------------
// Chemist order.  Block (cs, d): columns c in [cs C, cs C + C) of the fixed d;
// threads over f = a + n1 b of the contiguous plane ERI(:,:,c,d).  J(a,b) sums
// over the block's c in registers (one atomic per f and block); K(a,d) is the
// block's sum over b and c: per-thread sums meet in LDS, one atomic per a.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_jk_na_chem(const double *__restrict__ E, const double *__restrict__ P,
                                                   double *__restrict__ J, double *__restrict__ K, int n1, int nb,
                                                   int C) {
    extern __shared__ double kacc[];     // [n1]
    const int d = blockIdx.y, c0 = blockIdx.x * C, c1 = min(nb, c0 + C);
    const size_t plane = (size_t)n1 * nb;
    for (int i = threadIdx.x; i < n1; i += blockDim.x) kacc[i] = 0.0;
    __syncthreads();
    const double *Ed = E + plane * ((size_t)nb * d);
    for (size_t f = threadIdx.x; f < plane; f += blockDim.x) {
        const int a = (int)(f % n1), b = (int)(f / n1);
        double jv = 0.0, kv = 0.0;
        for (int c = c0; c < c1; ++c) {
            const double e = Ed[f + plane * c];
            jv += e * P[c + (size_t)nb * d];
            kv += e * P[b + (size_t)nb * c];
        }
        if (jv != 0.0) atomicAdd(&J[a + (size_t)n1 * b], jv);
        if (kv != 0.0) atomicAdd(&kacc[a], kv);
    }
    __syncthreads();
    for (int a = threadIdx.x; a < n1; a += blockDim.x)
        if (kacc[a] != 0.0) atomicAdd(&K[a + (size_t)n1 * d], -0.5 * kacc[a]);
}

// ---------------------------------------------------------------------------
// Physicist order.  Block (cs, d) as above; threads over (a, c) of the block,
// each summing over b: J(a,c) in registers (one atomic per task), K(a,d)
// through LDS as in the chemist kernel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_jk_na_phys(const double *__restrict__ E, const double *__restrict__ P,
                                                   double *__restrict__ J, double *__restrict__ K, int n1, int nb,
                                                   int C) {
    extern __shared__ double kacc[];     // [n1]
    const int d = blockIdx.y, c0 = blockIdx.x * C, c1 = min(nb, c0 + C);
    const size_t plane = (size_t)n1 * nb;
    for (int i = threadIdx.x; i < n1; i += blockDim.x) kacc[i] = 0.0;
    __syncthreads();
    const int ntask = n1 * (c1 - c0);
    for (int task = threadIdx.x; task < ntask; task += blockDim.x) {
        const int a = task % n1, c = c0 + task / n1;
        const double *Ec = E + a + plane * (c + (size_t)nb * d);
        double jv = 0.0, kv = 0.0;
        for (int b = 0; b < nb; ++b) {
            const double e = Ec[(size_t)n1 * b];
            jv += e * P[b + (size_t)nb * d];
            kv += e * P[c + (size_t)nb * b];
        }
        if (jv != 0.0) atomicAdd(&J[a + (size_t)n1 * c], jv);
        if (kv != 0.0) atomicAdd(&kacc[a], kv);
    }
    __syncthreads();
    for (int a = threadIdx.x; a < n1; a += blockDim.x)
        if (kacc[a] != 0.0) atomicAdd(&K[a + (size_t)n1 * d], -0.5 * kacc[a]);
}

// The resident tensor's layout and the engine's stream, from tcg_host.cpp.
int ctx_resident_eri(const tcg_ctx *c, const double **eri, size_t *count, int *n1, int *physicist, int *nbas,
                     hipStream_t *st);
void ctx_set_error(tcg_ctx *c, const char *msg);

}  // namespace tcg

using namespace tcg;

extern "C" int tcg_jk_stored_noalloc(tcg_ctx *c, int n1, int nbas, const double *P, double *J, double *K) {
    if (!c) return TCG_E_ARG;
    char msg[256];
    if (!P || !J || !K || nbas <= 0 || n1 <= 0 || n1 > nbas) {
        ctx_set_error(c, "tcg_jk_stored_noalloc: bad arguments");
        return TCG_E_ARG;
    }
    const size_t nn = (size_t)nbas * nbas, nj = (size_t)n1 * nbas;
    if (nn + 2 * nj > kScratchDoubles) {
        snprintf(msg, sizeof msg,
                 "tcg_jk_stored_noalloc: nbas %d, n1 %d need %zu doubles of scratch, more than its %zu "
                 "(nbas^2 + 2 n1 nbas <= %zu)",
                 nbas, n1, nn + 2 * nj, kScratchDoubles, kScratchDoubles);
        ctx_set_error(c, msg);
        return TCG_E_ARG;
    }
    const double *eri = nullptr;
    size_t count = 0;
    int rn1 = 0, phys = 0, rnb = 0;
    hipStream_t st = nullptr;
    if (ctx_resident_eri(c, &eri, &count, &rn1, &phys, &rnb, &st) != 0 || !eri || rn1 != n1 || rnb != nbas ||
        count != (size_t)n1 * nbas * nbas * nbas) {
        ctx_set_error(c, "tcg_jk_stored_noalloc: no resident ERI of this shape (call tcg_eri_build)");
        return TCG_E_STATE;
    }
    double *base = scratch_ptr();
    if (!base) {
        ctx_set_error(c, "tcg_jk_stored_noalloc: scratch not resolved on the current device (tcg_create "
                         "was called on another device)");
        return TCG_E_STATE;
    }
    double *dP = base, *dJ = base + nn, *dK = dJ + nj;
    auto chk = [&](hipError_t e, const char *what) -> int {
        if (e == hipSuccess) return 0;
        snprintf(msg, sizeof msg, "tcg_jk_stored_noalloc: %s: %s", what, hipGetErrorString(e));
        ctx_set_error(c, msg);
        return TCG_E_DEVICE;
    };
    int rc;
    if ((rc = chk(hipMemcpyAsync(dP, P, sizeof(double) * nn, hipMemcpyHostToDevice, st), "upload P"))) return rc;
    if ((rc = chk(launch_fill(dJ, 2 * nj, 0.0, st), "clear J, K"))) return rc;
    // columns per block: enough blocks to fill the chip, each streaming a few
    // hundred KB; the physicist kernel's tasks are (a, c) pairs of the block
    const int C = phys ? std::max(1, std::min(nbas, 2048 / n1)) : std::min(nbas, 64);
    const dim3 grid((unsigned)((nbas + C - 1) / C), (unsigned)nbas);
    const size_t lds = sizeof(double) * (size_t)n1;
    if (phys)
        hipLaunchKernelGGL(k_jk_na_phys, grid, dim3(256), lds, st, eri, dP, dJ, dK, n1, nbas, C);
    else
        hipLaunchKernelGGL(k_jk_na_chem, grid, dim3(256), lds, st, eri, dP, dJ, dK, n1, nbas, C);
    if ((rc = chk(hipGetLastError(), "J/K kernel"))) return rc;
    // J, K (nbas, nbas) column-major: rows < n1 from the scratch rows, the rest zero
    std::memset(J, 0, sizeof(double) * nn);
    std::memset(K, 0, sizeof(double) * nn);
    const size_t w = sizeof(double) * (size_t)n1, hp = sizeof(double) * (size_t)nbas;
    if ((rc = chk(hipMemcpy2DAsync(J, hp, dJ, w, w, (size_t)nbas, hipMemcpyDeviceToHost, st), "copy J"))) return rc;
    if ((rc = chk(hipMemcpy2DAsync(K, hp, dK, w, w, (size_t)nbas, hipMemcpyDeviceToHost, st), "copy K"))) return rc;
    return chk(hipStreamSynchronize(st), "synchronize");
}
