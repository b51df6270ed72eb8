// DCT Poisson solve as dense transforms on MFMA (CDNA4 / gfx950).
//
// Math: src/numeric/cpu/cpu_poisson.cpp:3-40 (FFTW REDFT10/REDFT01/RODFT01
// conventions, spectral divide, index shift for the sine transforms).  Here
// every axis transform is a matrix product on the matrix cores:
// v_mfma_f32_16x16x4_f32 (FP32 path) or v_mfma_f64_16x16x4_f64 (FP64 path),
// LDS-staged 64-wide wavefront tiles.
#include <cmath>
#include <vector>

#include "poisson.h"

namespace sphip {
namespace {

template <typename T>
using vec4 = T __attribute__((ext_vector_type(4)));

template <typename T>
__device__ inline vec4<T> mfma16x16x4(T a, T b, vec4<T> c);

template <>
__device__ inline vec4<float> mfma16x16x4<float>(float a, float b, vec4<float> c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <>
__device__ inline vec4<double> mfma16x16x4<double>(double a, double b, vec4<double> c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename T>
struct GemmBatch {
    const T* A[4];
    const T* B[4];
    T* C[4];
};

enum { kEpiNone = 0, kEpiSpectral = 1 };

template <typename T>
struct SpectralEpi {
    const T* u;   // u_j, length >= M
    const T* v;   // v_k, length >= N
    int nx, ny;   // unpadded sizes: rows >= nx or cols >= ny are written as 0
    T invNorm;    // 1 / (4 nx ny)
};

// C[M x N] = A[M x K] * B[K x N], all row-major.  Block = 4 waves in a 2x2
// arrangement; each wave owns (16*TM) x (16*TN) of the block tile and issues
// TM*TN 16x16x4 MFMAs per k-step.  The MFMA operand layout: lane l supplies
// A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16]; accumulator element r of
// lane l is C[4 * (l / 16) + r][l % 16] for f32 and C[4 * r + l / 16][l % 16]
// for f64 (both measured on gfx950).
template <typename T, int TM, int TN, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmBatch<T> batch, int M, int N, int K, int lda,
                                              int ldb, int ldc, SpectralEpi<T> epi,
                                              const int* skip) {
    if (skip && *skip) return;   // device loop finished: uniform early exit
    constexpr int BM = 32 * TM, BN = 32 * TN;
    constexpr int BK = sizeof(T) == 4 ? 32 : 16;
    // A is kept row-major in LDS (k fastest).  The MFMA read has 16 lanes on
    // 16 rows at one k and the next 16 lanes at k+1: a row stride of 2 mod 32
    // banks (FP32) or an odd number of doubles (FP64) makes it conflict-free.
    constexpr int LDA_S = sizeof(T) == 4 ? BK + 2 : BK + 1;
    // B is k-major with rows 16 banks apart: lanes read consecutive n.
    constexpr int LDB_S = sizeof(T) == 4 ? BN + 16 : BN + 8;
    constexpr int VEC = 16 / sizeof(T);                 // elements per 16-byte load
    constexpr int A_LD = BM * BK / (256 * VEC);          // 16-byte loads per thread
    constexpr int B_LD = BK * BN / (256 * VEC);
    static_assert(A_LD >= 1 && B_LD >= 1, "tile too small for 256 threads");
    __shared__ __attribute__((aligned(16))) T As[2][BM][LDA_S];
    __shared__ __attribute__((aligned(16))) T Bs[2][BK][LDB_S];
    using V = T __attribute__((ext_vector_type(16 / sizeof(T))));

    const T* __restrict__ A = batch.A[blockIdx.z];
    const T* __restrict__ B = batch.B[blockIdx.z];
    T* __restrict__ C = batch.C[blockIdx.z];
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lrow = lane & 15, lk = lane >> 4;

    // FP32: each BK-long chunk is accumulated by the MFMAs in FP32 and then
    // added into an FP64 accumulator, so the rounding grows with BK, not K.
    constexpr bool kChunked = sizeof(T) == 4;
    vec4<T> acc[TM][TN];
    vec4<double> acc64[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            acc[i][j] = vec4<T>{0, 0, 0, 0};
            acc64[i][j] = vec4<double>{0, 0, 0, 0};
        }

    V ra[A_LD], rb[B_LD];
    auto loadTiles = [&](int k0) {
#pragma unroll
        for (int q = 0; q < A_LD; ++q) {
            const int e = tid + 256 * q;
            const int r = e / (BK / VEC), c = (e % (BK / VEC)) * VEC;
            ra[q] = *reinterpret_cast<const V*>(A + static_cast<long>(m0 + r) * lda + k0 + c);
        }
#pragma unroll
        for (int q = 0; q < B_LD; ++q) {
            const int e = tid + 256 * q;
            const int r = e / (BN / VEC), c = (e % (BN / VEC)) * VEC;
            rb[q] = *reinterpret_cast<const V*>(B + static_cast<long>(k0 + r) * ldb + n0 + c);
        }
    };
    auto storeTiles = [&](int buf) {
#pragma unroll
        for (int q = 0; q < A_LD; ++q) {
            const int e = tid + 256 * q;
            const int r = e / (BK / VEC), c = (e % (BK / VEC)) * VEC;
#pragma unroll
            for (int v = 0; v < VEC; ++v) As[buf][r][c + v] = ra[q][v];
        }
#pragma unroll
        for (int q = 0; q < B_LD; ++q) {
            const int e = tid + 256 * q;
            const int r = e / (BN / VEC), c = (e % (BN / VEC)) * VEC;
            *reinterpret_cast<V*>(&Bs[buf][r][c]) = rb[q];
        }
    };

    loadTiles(0);
    storeTiles(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += BK) {
        const bool more = k0 + BK < K;
        if (more) loadTiles(k0 + BK);   // in flight while the MFMAs run
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[buf][wm * 16 * TM + i * 16 + lrow][kk + lk];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = Bs[buf][kk + lk][wn * 16 * TN + j * 16 + lrow];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = mfma16x16x4<T>(a[i], b[j], acc[i][j]);
        }
        if (kChunked) {
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc64[i][j][r] += static_cast<double>(acc[i][j][r]);
                    acc[i][j] = vec4<T>{0, 0, 0, 0};
                }
        }
        if (more) storeTiles(buf ^ 1);   // the other buffer: no reader is using it
        __syncthreads();
        buf ^= 1;
    }

#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = n0 + wn * 16 * TN + j * 16 + lrow;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // accumulator row of element r: f32 4*(l/16) + r, f64 4*r + l/16
                const int rib = sizeof(T) == 8 ? 4 * r + lk : 4 * lk + r;
                const int row = m0 + wm * 16 * TM + i * 16 + rib;
                const double full = kChunked ? acc64[i][j][r] : static_cast<double>(acc[i][j][r]);
                T val = static_cast<T>(full);
                if (EPI == kEpiSpectral) {
                    // P = F / (4 nx ny) / (u_j^2 + v_k^2), P[0][0] = 0 (cpu_poisson.cpp:163-173)
                    if (row < epi.nx && col < epi.ny) {
                        const T uj = epi.u[row], vk = epi.v[col];
                        const T den = uj * uj + vk * vk;
                        val = den > T(0) ? static_cast<T>(full * static_cast<double>(epi.invNorm) /
                                                          static_cast<double>(den))
                                         : T(0);
                    } else {
                        val = T(0);
                    }
                }
                C[static_cast<long>(row) * ldc + col] = val;
            }
        }
}

template <typename T, int EPI>
void gemmDispatch(int batch, const GemmBatch<T>& gb, int M, int N, int K, int lda, int ldb,
                  int ldc, const SpectralEpi<T>& epi, hipStream_t s, const int* skip = nullptr) {
    if (M % 64 || N % 64 || K % 64) throw std::runtime_error("gemm: dims must be multiples of 64");
    if (batch < 1 || batch > 4) throw std::runtime_error("gemm: batch must be 1..4");
    const long tiles64 = static_cast<long>(M / 64) * (N / 64) * batch;
    if (tiles64 >= 256) {
        dim3 grid(N / 64, M / 64, batch);
        launchKernel(k_gemm<T, 2, 2, EPI>, grid, 256, 0, s, gb, M, N, K, lda, ldb, ldc, epi, skip);
    } else {
        dim3 grid(N / 32, M / 32, batch);
        launchKernel(k_gemm<T, 1, 1, EPI>, grid, 256, 0, s, gb, M, N, K, lda, ldb, ldc, epi, skip);
    }
    SPHIP_CHECK(hipGetLastError());
}

template <typename T>
T* devAlloc(size_t n) {
    void* p = nullptr;
    SPHIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T*>(p);
}

template <typename T>
T* upload(const std::vector<double>& h, hipStream_t s) {
    std::vector<T> t(h.begin(), h.end());
    T* d = devAlloc<T>(t.size());
    SPHIP_CHECK(hipMemcpyAsync(d, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice, s));
    SPHIP_CHECK(hipStreamSynchronize(s));
    return d;
}

}  // namespace

template <typename T>
void launchGemm(int batch, const T* const* A, const T* const* B, T* const* C, int M, int N,
                int K, int lda, int ldb, int ldc, hipStream_t stream) {
    GemmBatch<T> gb{};
    for (int i = 0; i < batch && i < 4; ++i) {
        gb.A[i] = A[i];
        gb.B[i] = B[i];
        gb.C[i] = C[i];
    }
    gemmDispatch<T, kEpiNone>(batch, gb, M, N, K, lda, ldb, ldc, SpectralEpi<T>{}, stream);
}

template <typename T>
PoissonPlan<T>::PoissonPlan(int nx, int ny, double stepX, double stepY, hipStream_t stream)
    : nx_(nx), ny_(ny), mx_(roundUp(nx, 64)), my_(roundUp(ny, 64)), stream_(stream) {
    if (nx < 2 || ny < 2) throw std::invalid_argument("Poisson grid must be at least 2x2");
    if (!(stepX > 0) || !(stepY > 0)) throw std::invalid_argument("bin step must be positive");
    const double pi = 3.14159265358979323846;
    const size_t gx = static_cast<size_t>(mx_) * my_;
    // Transform matrices in FP64, zero in the padding (cpu_poisson.cpp:14-36).
    std::vector<double> cyT(static_cast<size_t>(my_) * my_, 0.0), ayT(cyT.size(), 0.0),
        bsv(cyT.size(), 0.0);
    std::vector<double> cx(static_cast<size_t>(mx_) * mx_, 0.0), sxu(cx.size(), 0.0),
        ax(cx.size(), 0.0);
    std::vector<double> u(mx_, 0.0), v(my_, 0.0);
    for (int j = 0; j < nx; ++j) u[j] = pi * j / (nx * stepX);
    for (int k = 0; k < ny; ++k) v[k] = pi * k / (ny * stepY);
    for (int y = 0; y < ny; ++y)
        for (int k = 0; k < ny; ++k) {
            // forward REDFT10 along y: T0[x][k] = sum_y rho[x][y] * 2 cos(pi k (y+1/2)/ny)
            cyT[static_cast<size_t>(y) * my_ + k] = 2.0 * std::cos(pi * k * (y + 0.5) / ny);
            // inverse REDFT01 along y: T1[j][y] = sum_k P[j][k] c_k cos(pi k (y+1/2)/ny)
            ayT[static_cast<size_t>(k) * my_ + y] =
                (k == 0 ? 1.0 : 2.0) * std::cos(pi * k * (y + 0.5) / ny);
            // RODFT01 along y of P shifted left and scaled by v (cpu_poisson.cpp:197-206):
            // T2[j][y] = sum_{k>=1} P[j][k] 2 v_k sin(pi k (y+1/2)/ny)
            bsv[static_cast<size_t>(k) * my_ + y] =
                k == 0 ? 0.0 : 2.0 * v[k] * std::sin(pi * k * (y + 0.5) / ny);
        }
    for (int x = 0; x < nx; ++x)
        for (int j = 0; j < nx; ++j) {
            cx[static_cast<size_t>(j) * mx_ + x] = 2.0 * std::cos(pi * j * (x + 0.5) / nx);
            ax[static_cast<size_t>(x) * mx_ + j] =
                (j == 0 ? 1.0 : 2.0) * std::cos(pi * j * (x + 0.5) / nx);
            // RODFT01 along x of P shifted down and scaled by u (cpu_poisson.cpp:181-195)
            sxu[static_cast<size_t>(x) * mx_ + j] =
                j == 0 ? 0.0 : 2.0 * u[j] * std::sin(pi * j * (x + 0.5) / nx);
        }
    cyT_ = upload<T>(cyT, stream);
    ayT_ = upload<T>(ayT, stream);
    bsv_ = upload<T>(bsv, stream);
    cx_ = upload<T>(cx, stream);
    sxu_ = upload<T>(sxu, stream);
    ax_ = upload<T>(ax, stream);
    // the reference computes u, v in float (cpu_poisson.cpp:88-94); the FP32
    // path keeps them as float, the FP64 path in double
    std::vector<double> uf(u), vf(v);
    if (sizeof(T) == 4) {
        for (int j = 0; j < nx; ++j)
            uf[j] = static_cast<float>(3.14159265358979323846f) * static_cast<float>(j) /
                    (static_cast<float>(nx) * static_cast<float>(stepX));
        for (int k = 0; k < ny; ++k)
            vf[k] = static_cast<float>(3.14159265358979323846f) * static_cast<float>(k) /
                    (static_cast<float>(ny) * static_cast<float>(stepY));
    }
    u_ = upload<T>(uf, stream);
    v_ = upload<T>(vf, stream);
    invNorm_ = static_cast<T>(1.0 / (4.0 * nx * ny));
    for (T** b : {&rho_, &phi_, &ex_, &ey_, &t0_, &p_, &t1_, &t2_}) {
        *b = devAlloc<T>(gx);
        SPHIP_CHECK(hipMemsetAsync(*b, 0, gx * sizeof(T), stream));
    }
    SPHIP_CHECK(hipStreamSynchronize(stream));
}

template <typename T>
PoissonPlan<T>::~PoissonPlan() {
    for (T* b : {rho_, phi_, ex_, ey_, t0_, p_, t1_, t2_, cyT_, cx_, ayT_, bsv_, sxu_, ax_, u_, v_})
        if (b) (void)hipFree(b);
}

template <typename T>
double PoissonPlan<T>::flops(bool withPhi) const {
    const double sq_y = 2.0 * mx_ * my_ * my_, sq_x = 2.0 * mx_ * mx_ * my_;
    // forward: y then x; inverse: 2 products along y, 2 (or 3) along x
    return sq_y + sq_x + 2 * sq_y + (withPhi ? 3 : 2) * sq_x;
}

template <typename T>
void PoissonPlan<T>::solve(bool withPhi, const int* skip) {
    const int M = mx_, N = my_;
    SpectralEpi<T> none{};
    {   // T0 = rho * CyT
        GemmBatch<T> gb{};
        gb.A[0] = rho_; gb.B[0] = cyT_; gb.C[0] = t0_;
        gemmDispatch<T, kEpiNone>(1, gb, M, N, N, N, N, N, none, stream_, skip);
    }
    {   // P = spectral(Cx * T0)
        GemmBatch<T> gb{};
        gb.A[0] = cx_; gb.B[0] = t0_; gb.C[0] = p_;
        SpectralEpi<T> epi{u_, v_, nx_, ny_, invNorm_};
        gemmDispatch<T, kEpiSpectral>(1, gb, M, N, M, M, N, N, epi, stream_, skip);
    }
    {   // T1 = P * AyT, T2 = P * Bsv
        GemmBatch<T> gb{};
        gb.A[0] = p_; gb.B[0] = ayT_; gb.C[0] = t1_;
        gb.A[1] = p_; gb.B[1] = bsv_; gb.C[1] = t2_;
        gemmDispatch<T, kEpiNone>(2, gb, M, N, N, N, N, N, none, stream_, skip);
    }
    {   // xi_x = Sxu * T1, xi_y = Ax * T2, phi = Ax * T1
        GemmBatch<T> gb{};
        gb.A[0] = sxu_; gb.B[0] = t1_; gb.C[0] = ex_;
        gb.A[1] = ax_;  gb.B[1] = t2_; gb.C[1] = ey_;
        gb.A[2] = ax_;  gb.B[2] = t1_; gb.C[2] = phi_;
        gemmDispatch<T, kEpiNone>(withPhi ? 3 : 2, gb, M, N, M, M, N, N, none, stream_, skip);
    }
}

template class PoissonPlan<float>;
template class PoissonPlan<double>;
template void launchGemm<float>(int, const float* const*, const float* const*, float* const*,
                                int, int, int, int, int, int, hipStream_t);
template void launchGemm<double>(int, const double* const*, const double* const*,
                                 double* const*, int, int, int, int, int, int, hipStream_t);

}  // namespace sphip
