// The dense algebra of the k-space (Bloch, "SAO") RHF loop on the GPU:
//
//   transform_AO_to_SAO / transform_SAO_to_AO      src/Tools/SAO_Transfer.f90:1-233
//   get_X_from_overlap_complex_blocks            src/Energies/Energies_tools.f90:328-380
//     (+ canonical_X_complex :471, shift_dropped_complex :514)
//   diagonalize_block_diagonal_complex_para      src/Tools/general_tools.f90:309-339
//     (+ clean_and_symmetrize, src/Tools/SAO_Transfer.f90:235-257)
//   the error vector, energy sums, occupations and density of RHF_SAO
//                                                src/Energies/SAO/RHF_SAO.f90:237-375
//
// Layouts.  Cells and k points share one index, x fastest (k = kx + nx (ky + ny kz)).
// A k block M_k is nfuc x nfuc complex (double2), column-major, blocks
// consecutive: M[k nf^2 + a + nf b].  ph[k N + R] = exp(+i k.R) (the reference's
// phase(k,R), built on the host with the same reduction of the angle).
// AO matrices are nbas x nbas real, column-major; AO index = a + nf R.
//
// Every per-k kernel runs one workgroup per k block with the block's matrices
// in LDS (three n x n complex buffers: n <= kMaxBlock).  The eigensolver is a
// two-sided cyclic Jacobi method for Hermitian matrices (round-robin ordering,
// n/2 disjoint rotations per step, all applied at once): each rotation is the
// real Jacobi rotation of the phase-rotated 2x2 block, V = D R D^H with
// D = diag(1, exp(-i phi)), phi = arg A(p,q).
#include <hip/hip_runtime.h>

#include <cmath>

#include "tcg_device.hpp"

namespace tcg {

namespace {

__device__ inline double2 c_mul(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// conj(a) b
__device__ inline double2 c_mulc(double2 a, double2 b) { return double2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }
__device__ inline double c_abs(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }

constexpr double kLindepShift = 1e4;    // shift_dropped_complex (Energies_tools.f90:521)
constexpr double kCleanTol = 1e-15;     // clean_and_symmetrize (SAO_Transfer.f90:243)
constexpr double kSkipTol = 1e-15;      // transform_SAO_to_AO skips |M_k(a,b)| below this (:194)
constexpr double kLindep = 1e-6;        // get_X_from_overlap_complex_blocks (Energies_tools.f90:343)
constexpr int kMaxSweeps = 40;
constexpr int kOccLds = 8192;          // levels staged in LDS by k_occupy (64 KiB)

// block sum over all threads of the workgroup (blockDim a multiple of 64, <= 1024)
__device__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}

__device__ double block_max(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s = fmax(s, red[i]);
    return s;
}

// C = op(A) B for n x n blocks in LDS (op: conj-transpose when ha)
__device__ void lds_gemm(const double2 *A, const double2 *B, double2 *C, int n, bool ha) {
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e % n, j = e / n;
        double2 s{0.0, 0.0};
        if (ha) {
            for (int l = 0; l < n; ++l) {
                const double2 t = c_mulc(A[l + n * i], B[l + n * j]);
                s.x += t.x;
                s.y += t.y;
            }
        } else {
            for (int l = 0; l < n; ++l) {
                const double2 t = c_mul(A[i + n * l], B[l + n * j]);
                s.x += t.x;
                s.y += t.y;
            }
        }
        C[e] = s;
    }
}

__device__ void lds_load(double2 *dst, const double2 *src, int n) {
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) dst[e] = src[e];
}

__device__ void lds_identity(double2 *Z, int n) {
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) Z[e] = double2{(e % n) == (e / n) ? 1.0 : 0.0, 0.0};
}

// Cyclic Jacobi on the Hermitian M (n x n, LDS, both triangles kept); the
// rotations are accumulated into Z (Z <- Z V).  On return M is diagonal to
// the stopping tolerance: off-diagonal Frobenius norm <= 1e-15 of the
// Frobenius norm of M.  rot: LDS scratch of 3 doubles per pair, red: 16 doubles.
__device__ void jacobi(double2 *M, double2 *Z, int n, double *rot, int *pairs, double *red) {
    double nrm = 0.0;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) nrm += M[e].x * M[e].x + M[e].y * M[e].y;
    nrm = block_sum(nrm, red);
    const double tol2 = 1e-30 * nrm;
    const int m = n + (n & 1);              // players, a dummy index n when n is odd
    const int np = m / 2;
    for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
        double off = 0.0;
        for (int e = threadIdx.x; e < n * n; e += blockDim.x)
            if ((e % n) != (e / n)) off += M[e].x * M[e].x + M[e].y * M[e].y;
        off = block_sum(off, red);
        if (!(off > tol2)) break;
        for (int r = 0; r < m - 1; ++r) {
            // round-robin pairs: (r, m-1) and ((r+s) mod (m-1), (r-s) mod (m-1)), s = 1 .. np-1
            for (int s = threadIdx.x; s < np; s += blockDim.x) {
                int p, q;
                if (s == 0) {
                    p = r;
                    q = m - 1;
                } else {
                    p = (r + s) % (m - 1);
                    q = (r - s + (m - 1)) % (m - 1);
                }
                if (p > q) {
                    const int t = p;
                    p = q;
                    q = t;
                }
                double c = 1.0, sr = 0.0, si = 0.0;
                if (q < n) {
                    const double2 b = M[p + n * q];
                    const double ab = c_abs(b);
                    if (ab > 1e-300) {
                        const double al = M[p + n * p].x, ga = M[q + n * q].x;
                        const double th = (ga - al) / (2.0 * ab);
                        double t;
                        if (fabs(th) > 1e150) t = 0.5 / th;
                        else t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        const double s_ = t * c;
                        sr = s_ * b.x / ab;         // sigma = s exp(i phi)
                        si = s_ * b.y / ab;
                    }
                }
                pairs[2 * s] = p;
                pairs[2 * s + 1] = q;
                rot[3 * s] = c;
                rot[3 * s + 1] = sr;
                rot[3 * s + 2] = si;
            }
            __syncthreads();
            // columns: M <- M V, Z <- Z V.  V = [[c, sigma], [-conj(sigma), c]] on (p, q)
            for (int w = threadIdx.x; w < np * n; w += blockDim.x) {
                const int s = w / n, i = w % n;
                const int p = pairs[2 * s], q = pairs[2 * s + 1];
                if (q >= n) continue;
                const double c = rot[3 * s];
                const double2 sg{rot[3 * s + 1], rot[3 * s + 2]};
                if (sg.x == 0.0 && sg.y == 0.0) continue;
                const double2 sgc{sg.x, -sg.y};
                {
                    const double2 x = M[i + n * p], y = M[i + n * q];
                    const double2 a1 = c_mul(sgc, y), a2 = c_mul(sg, x);
                    M[i + n * p] = double2{c * x.x - a1.x, c * x.y - a1.y};
                    M[i + n * q] = double2{a2.x + c * y.x, a2.y + c * y.y};
                }
                {
                    const double2 x = Z[i + n * p], y = Z[i + n * q];
                    const double2 a1 = c_mul(sgc, y), a2 = c_mul(sg, x);
                    Z[i + n * p] = double2{c * x.x - a1.x, c * x.y - a1.y};
                    Z[i + n * q] = double2{a2.x + c * y.x, a2.y + c * y.y};
                }
            }
            __syncthreads();
            // rows: M <- V^H M.  V^H = [[c, -sigma], [conj(sigma), c]]
            for (int w = threadIdx.x; w < np * n; w += blockDim.x) {
                const int s = w / n, j = w % n;
                const int p = pairs[2 * s], q = pairs[2 * s + 1];
                if (q >= n) continue;
                const double c = rot[3 * s];
                const double2 sg{rot[3 * s + 1], rot[3 * s + 2]};
                if (sg.x == 0.0 && sg.y == 0.0) continue;
                const double2 sgc{sg.x, -sg.y};
                const double2 x = M[p + n * j], y = M[q + n * j];
                const double2 a1 = c_mul(sg, y), a2 = c_mul(sgc, x);
                M[p + n * j] = double2{c * x.x - a1.x, c * x.y - a1.y};
                M[q + n * j] = double2{a2.x + c * y.x, a2.y + c * y.y};
            }
            __syncthreads();
            // the rotated pair: exact zeros off the diagonal, real diagonal
            for (int s = threadIdx.x; s < np; s += blockDim.x) {
                const int p = pairs[2 * s], q = pairs[2 * s + 1];
                if (q >= n || (rot[3 * s + 1] == 0.0 && rot[3 * s + 2] == 0.0)) continue;
                M[p + n * q] = double2{0.0, 0.0};
                M[q + n * p] = double2{0.0, 0.0};
                M[p + n * p].y = 0.0;
                M[q + n * q].y = 0.0;
            }
            __syncthreads();
        }
    }
}

// rank of diagonal element i in ascending order, ties by index (a stable sort)
__device__ int diag_rank(const double2 *M, int n, int i) {
    const double v = M[i + n * i].x;
    int r = 0;
    for (int j = 0; j < n; ++j) {
        const double u = M[j + n * j].x;
        r += (u < v || (u == v && j < i)) ? 1 : 0;
    }
    return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// Lattice Fourier transforms
// ---------------------------------------------------------------------------

// One workgroup per output k point (per lattice vector dR for the inverse):
// thread (ab, g) sums the lattice index over the residue class g mod G
// (G = blockDim / nf^2 groups), the G partial sums are added in LDS.
// out_k(a,b) = sum_R ph(k,R) sum_s f_s A_s(a, b + R nf)   (transform_AO_to_SAO on the
// unit-cell rows: element (a, c) of source s at p_s[a sa_s + c sc_s]).
__device__ inline double2 fwd_term_sum(const RowSources &src, const double2 *phk, int N, int nf, int ab, int g,
                                       int G) {
    const int a = ab % nf, b = ab / nf;
    double2 acc{0.0, 0.0};
    for (int R = g; R < N; R += G) {
        const long long col = b + (long long)nf * R;
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < src.n) v += src.f[q] * src.p[q][(long long)a * src.sa[q] + col * src.sc[q]];
        const double2 e = phk[R];
        acc.x += e.x * v;
        acc.y += e.y * v;
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_lattice_fwd(RowSources src, const double2 *__restrict__ ph, int N, int nf,
                                                     double2 *__restrict__ out) {
    __shared__ double2 part[256];
    const int nff = nf * nf, k = blockIdx.x, t = threadIdx.x;
    const double2 *phk = ph + (size_t)k * N;
    double2 *o = out + (size_t)k * nff;
    if (nff <= (int)blockDim.x) {
        const int G = blockDim.x / nff, ab = t % nff, g = t / nff;
        part[t] = g < G ? fwd_term_sum(src, phk, N, nf, ab, g, G) : double2{0.0, 0.0};
        __syncthreads();
        if (t < nff) {
            double2 sm{0.0, 0.0};
            for (int q = 0; q < G; ++q) {
                sm.x += part[t + nff * q].x;
                sm.y += part[t + nff * q].y;
            }
            o[t] = sm;
        }
    } else {
        for (int ab = t; ab < nff; ab += blockDim.x) o[ab] = fwd_term_sum(src, phk, N, nf, ab, 0, 1);
    }
}

// C[dR](a,b) = Re (1/N) sum_k ph(k,dR) M_k(a,b), elements |M_k(a,b)| < 1e-15 left
// out when skip (transform_SAO_to_AO, SAO_Transfer.f90:186-225).  maxabs (optional):
// max |C[dR]| into maxabs[dR].
__device__ inline double inv_term_sum(const double2 *Mk, const double2 *ph, int N, int nff, int dR, int skip, int ab,
                                      int g, int G) {
    double acc = 0.0;
    for (int k = g; k < N; k += G) {
        const double2 m = Mk[(size_t)k * nff + ab];
        if (skip && m.x * m.x + m.y * m.y < kSkipTol * kSkipTol) continue;
        const double2 e = ph[(size_t)k * N + dR];
        acc += e.x * m.x - e.y * m.y;
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_lattice_inv(const double2 *__restrict__ Mk, const double2 *__restrict__ ph,
                                                     int N, int nf, int skip, double *__restrict__ C,
                                                     double *__restrict__ maxabs) {
    __shared__ double part[256];
    __shared__ double red[16];
    const int nff = nf * nf, dR = blockIdx.x, t = threadIdx.x;
    double *o = C + (size_t)dR * nff;
    double mx = 0.0;
    if (nff <= (int)blockDim.x) {
        const int G = blockDim.x / nff, ab = t % nff, g = t / nff;
        part[t] = g < G ? inv_term_sum(Mk, ph, N, nff, dR, skip, ab, g, G) : 0.0;
        __syncthreads();
        if (t < nff) {
            double sm = 0.0;
            for (int q = 0; q < G; ++q) sm += part[t + nff * q];
            const double v = sm / N;
            o[t] = v;
            mx = fabs(v);
        }
    } else {
        for (int ab = t; ab < nff; ab += blockDim.x) {
            const double v = inv_term_sum(Mk, ph, N, nff, dR, skip, ab, 0, 1) / N;
            o[ab] = v;
            mx = fmax(mx, fabs(v));
        }
    }
    if (maxabs) {
        mx = block_max(mx, red);
        if (t == 0) maxabs[dR] = mx;
    }
}

// M(a + nf R, b + nf R') = C[R - R' (per direction, mod n)](a,b)   (nbas x nbas)
__global__ __launch_bounds__(256) void k_circulant(const double *__restrict__ C, int nf, int nx, int ny, int nz,
                                                   double *__restrict__ M) {
    const int N = nx * ny * nz, nb = nf * N;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nb * nb) return;
    const int i = (int)(idx % nb), j = (int)(idx / nb);
    const int R = i / nf, a = i % nf, Rp = j / nf, b = j % nf;
    const int rx = R % nx, ry = (R / nx) % ny, rz = R / (nx * ny);
    const int px = Rp % nx, py = (Rp / nx) % ny, pz = Rp / (nx * ny);
    const int dR = (rx - px + nx) % nx + nx * ((ry - py + ny) % ny + ny * ((rz - pz + nz) % nz));
    M[idx] = C[(size_t)dR * nf * nf + a + nf * b];
}

// B = A^T, n x n, 32 x 32 tiles through LDS
__global__ __launch_bounds__(256) void k_transpose(const double *__restrict__ A, int n, double *__restrict__ B) {
    __shared__ double t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;     // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int i = bx + tx, j = by + r;
        if (i < n && j < n) t[r][tx] = A[(size_t)i + (size_t)n * j];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = by + tx, j = bx + r;
        if (i < n && j < n) B[(size_t)i + (size_t)n * j] = t[tx][r];
    }
}

// Hrow(a, c) = A(a, c) + B(a, c), a < nf (unit-cell rows of T + V), ld nf
__global__ __launch_bounds__(256) void k_rows_sum(const double *__restrict__ A, const double *__restrict__ B, int nf,
                                                  int nb, double *__restrict__ H) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nf * nb) return;
    const int a = (int)(idx % nf), c = (int)(idx / nf);
    H[idx] = A[a + (size_t)nb * c] + B[a + (size_t)nb * c];
}

// ---------------------------------------------------------------------------
// Per-k blocks
// ---------------------------------------------------------------------------

// X_k from S_k (get_X_from_overlap_complex_blocks): eigen-decomposition of the
// block (upper triangle, as zheev 'U'); Lowdin S_k^-1/2 when every eigenvalue
// exceeds 1e-6, else canonical columns U(:,j)/sqrt(s_j) in ascending order with
// the dropped ones zero (dead[k n + j] = 1).  smin[k] = smallest eigenvalue.
__global__ void k_overlap_x(const double2 *__restrict__ Sk, int n, double2 *__restrict__ Xk,
                            int *__restrict__ dead, double *__restrict__ smin) {
    extern __shared__ double2 lds[];
    __shared__ double rot[3 * 32];
    __shared__ int pairs[2 * 32];
    __shared__ double red[16];
    __shared__ double sv[64];
    __shared__ int rk[64];
    __shared__ int ndrop;
    const int k = blockIdx.x, nn = n * n;
    double2 *M = lds, *Z = lds + nn, *W = lds + 2 * nn;
    lds_load(M, Sk + (size_t)k * nn, n);
    lds_identity(Z, n);
    if (threadIdx.x == 0) ndrop = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {      // upper triangle defines the matrix
        const int i = e % n, j = e / n;
        if (i > j) W[e] = double2{M[j + n * i].x, -M[j + n * i].y};
        else if (i == j) W[e] = double2{M[e].x, 0.0};
        else W[e] = M[e];
    }
    __syncthreads();
    lds_load(M, W, n);
    __syncthreads();
    jacobi(M, Z, n, rot, pairs, red);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int r = diag_rank(M, n, i);
        rk[i] = r;
        sv[i] = M[i + n * i].x;
        if (!(sv[i] > kLindep)) atomicAdd(&ndrop, 1);
        if (r == 0) smin[k] = sv[i];
    }
    __syncthreads();
    double2 *X = Xk + (size_t)k * nn;
    if (ndrop == 0) {
        // X = Z diag(s^-1/2) Z^H
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const int a = e % n, b = e / n;
            double2 s{0.0, 0.0};
            for (int l = 0; l < n; ++l) {
                const double d = 1.0 / sqrt(sv[l]);
                const double2 t = c_mul(Z[a + n * l], double2{Z[b + n * l].x, -Z[b + n * l].y});
                s.x += d * t.x;
                s.y += d * t.y;
            }
            X[e] = s;
        }
        for (int j = threadIdx.x; j < n; j += blockDim.x) dead[(size_t)k * n + j] = 0;
    } else {
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const int a = e % n, i = e / n;       // eigenvector i -> column rk[i]
            const bool keep = sv[i] > kLindep;
            const double d = keep ? 1.0 / sqrt(sv[i]) : 0.0;
            X[a + n * rk[i]] = keep ? double2{Z[a + n * i].x * d, Z[a + n * i].y * d} : double2{0.0, 0.0};
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x) dead[(size_t)k * n + rk[i]] = sv[i] > kLindep ? 0 : 1;
    }
}

// F'_k = X_k^H F_k X_k, dropped vectors shifted to 1e4, clean_and_symmetrize
// (lower triangle, |x| < 1e-15 -> 0, real diagonal), eigenpairs ascending:
// e[k n + r], C_k(:, r) = X_k U(:, r) (diagonalize_block_diagonal_complex_para and the
// back-transformation of RHF_SAO.f90:343-375).  X = nullptr: F_k itself (no
// transformation, no shift), eigenvectors into C.
__global__ void k_fock_diag(const double2 *__restrict__ Fk, const double2 *__restrict__ Xk,
                            const int *__restrict__ dead, int n, double *__restrict__ e, double2 *__restrict__ Ck) {
    extern __shared__ double2 lds[];
    __shared__ double rot[3 * 32];
    __shared__ int pairs[2 * 32];
    __shared__ double red[16];
    __shared__ int rk[64];
    const int k = blockIdx.x, nn = n * n;
    double2 *X = lds, *A = lds + nn, *M = lds + 2 * nn;
    if (Xk) {
        lds_load(M, Fk + (size_t)k * nn, n);
        lds_load(X, Xk + (size_t)k * nn, n);
        __syncthreads();
        lds_gemm(M, X, A, n, false);          // F X
        __syncthreads();
        lds_gemm(X, A, M, n, true);           // X^H (F X)
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += blockDim.x)
            if (dead[(size_t)k * n + j]) M[j + n * j] = double2{kLindepShift, 0.0};
    } else {
        lds_load(M, Fk + (size_t)k * nn, n);
    }
    __syncthreads();
    // clean_and_symmetrize: the lower triangle defines the matrix
    for (int e2 = threadIdx.x; e2 < nn; e2 += blockDim.x) {
        const int i = e2 % n, j = e2 / n;
        if (i > j) {
            double2 v = M[e2];
            if (c_abs(v) < kCleanTol) v = double2{0.0, 0.0};
            A[e2] = v;
        } else if (i == j) {
            A[e2] = double2{M[e2].x, 0.0};
        } else {
            double2 v = M[j + n * i];
            if (c_abs(v) < kCleanTol) v = double2{0.0, 0.0};
            A[e2] = double2{v.x, -v.y};
        }
    }
    __syncthreads();
    lds_load(M, A, n);
    lds_identity(A, n);
    __syncthreads();
    jacobi(M, A, n, rot, pairs, red);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int r = diag_rank(M, n, i);
        rk[i] = r;
        e[(size_t)k * n + r] = M[i + n * i].x;
    }
    __syncthreads();
    double2 *C = Ck + (size_t)k * nn;
    if (Xk) {
        for (int e2 = threadIdx.x; e2 < nn; e2 += blockDim.x) {
            const int a = e2 % n, i = e2 / n;
            double2 s{0.0, 0.0};
            for (int l = 0; l < n; ++l) {
                const double2 t = c_mul(X[a + n * l], A[l + n * i]);
                s.x += t.x;
                s.y += t.y;
            }
            C[a + n * rk[i]] = s;
        }
    } else {
        for (int e2 = threadIdx.x; e2 < nn; e2 += blockDim.x) {
            const int a = e2 % n, i = e2 / n;
            C[a + n * rk[i]] = A[e2];
        }
    }
}

// err_k = X_k^H (F_k P_k S_k - S_k P_k F_k) X_k   (RHF_SAO.f90:276-286 in k space;
// F, P, S Hermitian: S P F = (F P S)^H)
__global__ void k_scf_error(const double2 *__restrict__ Fk, const double2 *__restrict__ Pk,
                            const double2 *__restrict__ Sk, const double2 *__restrict__ Xk, int n,
                            double2 *__restrict__ err) {
    extern __shared__ double2 lds[];
    const int k = blockIdx.x, nn = n * n;
    const size_t o = (size_t)k * nn;
    double2 *B0 = lds, *B1 = lds + nn, *B2 = lds + 2 * nn;
    lds_load(B0, Pk + o, n);
    lds_load(B1, Sk + o, n);
    __syncthreads();
    lds_gemm(B0, B1, B2, n, false);           // P S
    __syncthreads();
    lds_load(B0, Fk + o, n);
    __syncthreads();
    lds_gemm(B0, B2, B1, n, false);           // F P S
    __syncthreads();
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e % n, j = e / n;
        const double2 x = B1[e], y = B1[j + n * i];
        B2[e] = double2{x.x - y.x, x.y + y.y};
    }
    __syncthreads();
    lds_load(B0, Xk + o, n);
    __syncthreads();
    lds_gemm(B2, B0, B1, n, false);           // E X
    __syncthreads();
    lds_gemm(B0, B1, B2, n, true);            // X^H E X
    __syncthreads();
    for (int e = threadIdx.x; e < nn; e += blockDim.x) err[o + e] = B2[e];
}

// P_k(a,b) = sum_i w[k n + i] C_k(a,i) conj(C_k(b,i))
__global__ __launch_bounds__(256) void k_density_k(const double2 *__restrict__ Ck, const double *__restrict__ w,
                                                   int N, int n, double2 *__restrict__ Pk) {
    const int nn = n * n;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * nn) return;
    const int k = idx / nn, r = idx % nn, a = r % n, b = r / n;
    const double2 *C = Ck + (size_t)k * nn;
    double2 s{0.0, 0.0};
    for (int i = 0; i < n; ++i) {
        const double wi = w[(size_t)k * n + i];
        if (wi == 0.0) continue;
        const double2 t = c_mul(C[a + n * i], double2{C[b + n * i].x, -C[b + n * i].y});
        s.x += wi * t.x;
        s.y += wi * t.y;
    }
    Pk[idx] = s;
}

// Occupations over the union of the k spectra (ne = N n values, one workgroup).
// frac = 0: weight 2 on the nocc lowest (ascending, ties by index: a stable
// sort; RHF_SAO.f90:352-395).  frac = 1: the guess rule of guess_occupy_cut
// (src/Energies/SAO/guess_Huckel_SAO.f90:141-192): the level at the cut, all
// values within tol = 1e-8 max(1, |e_cut|) of it, shares the remaining electrons
// evenly.  w[ne] weights, occ[ne] = 2 / 0 (frac = 0), sc[0] = HOMO-LUMO gap.
// The levels are staged in LDS when they fit (ne <= kOccLds).
__global__ __launch_bounds__(1024) void k_occupy(const double *__restrict__ e_g, int ne, int nocc, int frac,
                                                 double *__restrict__ w, int *__restrict__ occ,
                                                 double *__restrict__ sc) {
    extern __shared__ double e_l[];
    __shared__ double ecut, enext;
    __shared__ int nlo, nwin;
    if (threadIdx.x == 0) {
        ecut = 0.0;
        enext = 0.0;
        nlo = 0;
        nwin = 0;
    }
    const bool staged = ne <= kOccLds;
    if (staged)
        for (int i = threadIdx.x; i < ne; i += blockDim.x) e_l[i] = e_g[i];
    __syncthreads();
    const double *e = staged ? e_l : e_g;
    for (int i = threadIdx.x; i < ne; i += blockDim.x) {
        const double v = e[i];
        int r = 0;
        for (int j = 0; j < ne; ++j) {
            const double u = e[j];
            r += (u < v || (u == v && j < i)) ? 1 : 0;
        }
        if (r == nocc - 1) ecut = v;
        if (r == nocc) enext = v;
        if (!frac) {
            w[i] = r < nocc ? 2.0 : 0.0;
            occ[i] = r < nocc ? 2 : 0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) sc[0] = ne > nocc ? enext - ecut : 0.0;
    if (!frac) return;
    const double tol = 1e-8 * fmax(1.0, fabs(ecut));
    for (int i = threadIdx.x; i < ne; i += blockDim.x) {
        const double v = e[i];
        if (fabs(v - ecut) < tol) atomicAdd(&nwin, 1);
        else if (v < ecut) atomicAdd(&nlo, 1);
    }
    __syncthreads();
    const double share = 2.0 * (double)(nocc - nlo) / (double)nwin;
    for (int i = threadIdx.x; i < ne; i += blockDim.x) {
        const double v = e[i];
        double wi = 0.0;
        if (fabs(v - ecut) < tol) wi = share;
        else if (v < ecut) wi = 2.0;
        w[i] = wi;
        occ[i] = wi > 0.0 ? 2 : 0;
    }
}

// ---------------------------------------------------------------------------
// Energy sums, DIIS
// ---------------------------------------------------------------------------

// partial sums per workgroup: [0] sum P Tt, [1] sum P Vt over nbas^2 (Tt = T^T:
// sum_ij P(i,j) T(j,i)); [2] sum_{a<nf,c} J(a,c) P(c,a), [3] the same with K.
__global__ __launch_bounds__(256) void k_energy_parts(const double *__restrict__ P, const double *__restrict__ Tt,
                                                      const double *__restrict__ Vt, const double *__restrict__ J,
                                                      const double *__restrict__ K, int ldjk, int nf, int nb,
                                                      double *__restrict__ part) {
    __shared__ double red[16];
    const long long n2 = (long long)nb * nb, nr = (long long)nf * nb;
    const long long stride = (long long)gridDim.x * blockDim.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) {
        const double p = P[i];
        s0 += p * Tt[i];
        s1 += p * Vt[i];
    }
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += stride) {
        const int c = (int)(i % nb), a = (int)(i / nb);
        const double p = P[c + (size_t)nb * a];
        s2 += J[a + (size_t)ldjk * c] * p;
        s3 += K[a + (size_t)ldjk * c] * p;
    }
    s0 = block_sum(s0, red);
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    s3 = block_sum(s3, red);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = s0;
        part[4 * blockIdx.x + 1] = s1;
        part[4 * blockIdx.x + 2] = s2;
        part[4 * blockIdx.x + 3] = s3;
    }
}

// sc[0..3] = ET, EV, EJ, EK (EJ, EK scaled by 1/2 N_cell, RHF_SAO.f90:263-274), sc[4] =
// max of the nmax maxima (orbital error), in a fixed order.
__global__ __launch_bounds__(256) void k_energy_finish(const double *__restrict__ part, int nparts,
                                                       const double *__restrict__ maxes, int nmax, double ncell,
                                                       double *__restrict__ sc) {
    __shared__ double red[16];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nparts; i += blockDim.x)
        for (int q = 0; q < 4; ++q) s[q] += part[4 * i + q];
    double mx = 0.0;
    for (int i = threadIdx.x; i < nmax; i += blockDim.x) mx = fmax(mx, maxes[i]);
    for (int q = 0; q < 4; ++q) s[q] = block_sum(s[q], red);
    mx = block_max(mx, red);
    if (threadIdx.x == 0) {
        sc[0] = s[0];
        sc[1] = s[1];
        sc[2] = 0.5 * s[2] * ncell;
        sc[3] = 0.5 * s[3] * ncell;
        sc[4] = mx;
    }
}

// out[j] = Re <enew, hist[slot_j]> (Frobenius product over all k blocks), one workgroup per j
__global__ __launch_bounds__(256) void k_diis_dots(const double2 *__restrict__ enew, const double2 *__restrict__ hist,
                                                   long long len, DiisSlots sl, double *__restrict__ out) {
    __shared__ double red[16];
    const double2 *h = hist + (size_t)sl.slot[blockIdx.x] * len;
    double s = 0.0;
    for (long long i = threadIdx.x; i < len; i += blockDim.x) s += enew[i].x * h[i].x + enew[i].y * h[i].y;
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// out = sum_j w_j hist[slot_j]
__global__ __launch_bounds__(256) void k_diis_combine(const double2 *__restrict__ hist, long long len, DiisSlots sl,
                                                      double2 *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double2 s{0.0, 0.0};
    for (int j = 0; j < sl.m; ++j) {
        const double2 h = hist[(size_t)sl.slot[j] * len + i];
        s.x += sl.w[j] * h.x;
        s.y += sl.w[j] * h.y;
    }
    out[i] = s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned blocks_for(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

static size_t block_lds(int n) { return 3 * (size_t)n * n * sizeof(double2); }
static int block_threads(int n) { return n <= 16 ? 64 : 256; }

// dynamic LDS beyond 64 KiB needs the kernel's limit raised
template <class K>
static hipError_t allow_lds(K kern, size_t lds) {
    if (lds <= 65536) return hipSuccess;
    return hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_lattice_fwd(const RowSources &src, const double2 *ph, int N, int nf, double2 *out, hipStream_t st) {
    hipLaunchKernelGGL(k_lattice_fwd, dim3(N), dim3(256), 0, st, src, ph, N, nf, out);
    return hipGetLastError();
}

hipError_t launch_lattice_inv(const double2 *Mk, const double2 *ph, int N, int nf, int skip, double *C,
                              double *maxabs, int *nmax, hipStream_t st) {
    if (nmax) *nmax = N;
    hipLaunchKernelGGL(k_lattice_inv, dim3(N), dim3(256), 0, st, Mk, ph, N, nf, skip, C, maxabs);
    return hipGetLastError();
}

hipError_t launch_circulant(const double *C, int nf, const int ncl[3], double *M, hipStream_t st) {
    const long long nb = (long long)nf * ncl[0] * ncl[1] * ncl[2];
    hipLaunchKernelGGL(k_circulant, dim3(blocks_for(nb * nb)), dim3(256), 0, st, C, nf, ncl[0], ncl[1], ncl[2], M);
    return hipGetLastError();
}

hipError_t launch_transpose(const double *A, int n, double *B, hipStream_t st) {
    const unsigned t = (unsigned)((n + 31) / 32);
    hipLaunchKernelGGL(k_transpose, dim3(t, t), dim3(256), 0, st, A, n, B);
    return hipGetLastError();
}

hipError_t launch_rows_sum(const double *A, const double *B, int nf, int nb, double *H, hipStream_t st) {
    hipLaunchKernelGGL(k_rows_sum, dim3(blocks_for((long long)nf * nb)), dim3(256), 0, st, A, B, nf, nb, H);
    return hipGetLastError();
}

hipError_t launch_overlap_x(const double2 *Sk, int N, int n, double2 *Xk, int *dead, double *smin, hipStream_t st) {
    if (n < 1 || n > kMaxBlock) return hipErrorInvalidValue;
    const hipError_t ea = allow_lds(k_overlap_x, block_lds(n));
    if (ea != hipSuccess) return ea;
    hipLaunchKernelGGL(k_overlap_x, dim3(N), dim3(block_threads(n)), block_lds(n), st, Sk, n, Xk, dead, smin);
    return hipGetLastError();
}

hipError_t launch_fock_diag(const double2 *Fk, const double2 *Xk, const int *dead, int N, int n, double *e,
                            double2 *Ck, hipStream_t st) {
    if (n < 1 || n > kMaxBlock) return hipErrorInvalidValue;
    const hipError_t ea = allow_lds(k_fock_diag, block_lds(n));
    if (ea != hipSuccess) return ea;
    hipLaunchKernelGGL(k_fock_diag, dim3(N), dim3(block_threads(n)), block_lds(n), st, Fk, Xk, dead, n, e, Ck);
    return hipGetLastError();
}

hipError_t launch_scf_error(const double2 *Fk, const double2 *Pk, const double2 *Sk, const double2 *Xk, int N, int n,
                            double2 *err, hipStream_t st) {
    if (n < 1 || n > kMaxBlock) return hipErrorInvalidValue;
    const hipError_t ea = allow_lds(k_scf_error, block_lds(n));
    if (ea != hipSuccess) return ea;
    hipLaunchKernelGGL(k_scf_error, dim3(N), dim3(block_threads(n)), block_lds(n), st, Fk, Pk, Sk, Xk, n, err);
    return hipGetLastError();
}

hipError_t launch_density_k(const double2 *Ck, const double *w, int N, int n, double2 *Pk, hipStream_t st) {
    hipLaunchKernelGGL(k_density_k, dim3(blocks_for((long long)N * n * n)), dim3(256), 0, st, Ck, w, N, n, Pk);
    return hipGetLastError();
}

hipError_t launch_occupy(const double *e, int ne, int nocc, int frac, double *w, int *occ, double *sc,
                         hipStream_t st) {
    const size_t lds = ne <= kOccLds ? sizeof(double) * (size_t)ne : 0;
    hipLaunchKernelGGL(k_occupy, dim3(1), dim3(1024), lds, st, e, ne, nocc, frac, w, occ, sc);
    return hipGetLastError();
}

int energy_parts_blocks(int nb) { return (int)std::min<long long>(blocks_for((long long)nb * nb), 1024); }

hipError_t launch_energy(const double *P, const double *Tt, const double *Vt, const double *J, const double *K,
                         int ldjk, int nf, int nb, double *part, const double *maxes, int nmax, double ncell,
                         double *sc, hipStream_t st) {
    const int g = energy_parts_blocks(nb);
    hipLaunchKernelGGL(k_energy_parts, dim3(g), dim3(256), 0, st, P, Tt, Vt, J, K, ldjk, nf, nb, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_energy_finish, dim3(1), dim3(256), 0, st, part, g, maxes, nmax, ncell, sc);
    return hipGetLastError();
}

hipError_t launch_diis_dots(const double2 *enew, const double2 *hist, long long len, const DiisSlots &sl, double *out,
                            hipStream_t st) {
    if (sl.m < 1 || sl.m > kMaxDiis) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_diis_dots, dim3(sl.m), dim3(256), 0, st, enew, hist, len, sl, out);
    return hipGetLastError();
}

hipError_t launch_diis_combine(const double2 *hist, long long len, const DiisSlots &sl, double2 *out, hipStream_t st) {
    if (sl.m < 1 || sl.m > kMaxDiis) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_diis_combine, dim3(blocks_for(len)), dim3(256), 0, st, hist, len, sl, out);
    return hipGetLastError();
}

}  // namespace tcg
