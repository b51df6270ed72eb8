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

// s_waitcnt with only the vector-memory count limited (gfx9 encoding).
template <int N>
__device__ inline void waitVmem() {
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// One 16-byte copy per lane from global memory into LDS at dst + 16 * lane
// (dst wave-uniform): global_load_lds_dwordx4.
__device__ inline void copyToLds(const void* src, void* dst) {
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

// C = A * Bt^T: A is M x K and Bt is N x K, both row-major (k contiguous).
// One block per 32 x 32 tile of C, 4 waves in a 2 x 2 arrangement, each
// owning a 16 x 16 quadrant.  The block's operand panels (32 rows of A and
// of Bt, KP columns at a time) are copied straight from memory into LDS by
// global_load_lds_dwordx4 (each lane moves 16 bytes; a wave instruction
// fills 1 KB of a panel row), all issued at once.
// Panel rows are padded by 16 bytes, so the ds_read_b128 operand reads of 16
// rows hit distinct banks.  MFMA operand layout: lane l supplies A[l % 16][k]
// and B[k][l % 16] for the k of its row slot l / 16; lane (r, q) covers
// k = 16 c + 4 q + s in phase s of chunk c, the same k for both operands.
// Accumulator element e of lane l is C[4 (l / 16) + e][l % 16] for f32 and
// C[4 e + l / 16][l % 16] for f64 (both measured on gfx950).  FP32 phases
// accumulate at most 64 products in FP32 before they are added in FP64.
// TRANS stores C^T (row stride ldc) instead of C; element e of the stored
// matrix goes to C[e * cstep] (cstep 2 interleaves two outputs).
template <typename T>
struct PanelGeom {
    static constexpr int KP = sizeof(T) == 4 ? 256 : 128;     // panel columns (1 KB per row)
    static constexpr int LD = KP + 16 / static_cast<int>(sizeof(T));   // padded row
    static constexpr int ROWS = 32;
};

template <typename T, int EPI, bool TRANS>
__global__ __launch_bounds__(256) void k_gemm(GemmBatch<T> batch, int M, int N, int K, int lda,
                                              int ldb, int ldc, int cstep, SpectralEpi<T> epi,
                                              const int* skip) {
    if (skip && *skip) return;   // device loop finished: uniform early exit
    using G = PanelGeom<T>;
    constexpr int KP = G::KP, LD = G::LD;
    constexpr int VE = 16 / static_cast<int>(sizeof(T));   // elements per 16-byte vector
    constexpr int NV = 4 / VE;                              // vectors per lane per 16-k chunk
    using V = T __attribute__((ext_vector_type(VE)));
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    T* As = reinterpret_cast<T*>(smem);
    T* Bs = As + G::ROWS * LD;

    const T* __restrict__ A = batch.A[blockIdx.z];
    const T* __restrict__ Bt = batch.B[blockIdx.z];
    T* __restrict__ C = batch.C[blockIdx.z];
    const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int wm = wave >> 1, wn = wave & 1;

    double acc64[4] = {0.0, 0.0, 0.0, 0.0};
    constexpr int SEG = 1024 / static_cast<int>(sizeof(T));   // elements per 1 KB segment
    const T* arow = As + (wm * 16 + r) * LD + 4 * q;
    const T* brow = Bs + (wn * 16 + r) * LD + 4 * q;
    for (int kc = 0; kc < K; kc += KP) {
        const int kw = K - kc < KP ? K - kc : KP;   // multiple of 64
        if (kc > 0) __syncthreads();                 // previous panel fully read
        // wave w copies panel rows 8w .. 8w+7 of A and Bt in 1 KB segments
        // (the last one partial when kw * sizeof(T) is not a multiple of 1 KB)
        const int nseg = (kw + SEG - 1) / SEG;
        for (int i = 0; i < 8; ++i) {
            const int row = wave * 8 + i;
            for (int sg = 0; sg < nseg; ++sg) {
                const int col = sg * SEG;
                if (col + lane * VE < kw) {
                    copyToLds(A + static_cast<long>(m0 + row) * lda + kc + col + lane * VE,
                              As + row * LD + col);
                    copyToLds(Bt + static_cast<long>(n0 + row) * ldb + kc + col + lane * VE,
                              Bs + row * LD + col);
                }
            }
        }
        waitVmem<0>();
        __syncthreads();
        // FP32 accumulation over at most 256 k (64 products per element and
        // phase), then into FP64
        for (int k0 = 0; k0 < kw; k0 += 256) {
            const int k1 = kw - k0 < 256 ? kw : k0 + 256;
            vec4<T> acc[4];
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2) acc[s2] = vec4<T>{0, 0, 0, 0};
            // operands of the next 16 k are read from LDS while the MFMAs of
            // the current ones run (the last prefetch re-reads the same k)
            V av[NV], bv[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                av[v] = *reinterpret_cast<const V*>(arow + k0 + v * VE);
                bv[v] = *reinterpret_cast<const V*>(brow + k0 + v * VE);
            }
            for (int kb = k0; kb < k1; kb += 16) {
                const int kn = kb + 16 < k1 ? kb + 16 : kb;
                V an[NV], bn[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    an[v] = *reinterpret_cast<const V*>(arow + kn + v * VE);
                    bn[v] = *reinterpret_cast<const V*>(brow + kn + v * VE);
                }
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2)
                    acc[s2] = mfma16x16x4<T>(av[s2 / VE][s2 % VE], bv[s2 / VE][s2 % VE], acc[s2]);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    av[v] = an[v];
                    bv[v] = bn[v];
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
                acc64[e] += ((static_cast<double>(acc[0][e]) + static_cast<double>(acc[1][e])) +
                             static_cast<double>(acc[2][e])) + static_cast<double>(acc[3][e]);
        }
    }

#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int row = wm * 16 + (sizeof(T) == 8 ? 4 * e + q : 4 * q + e);
        const int col = wn * 16 + r;
        const int gr = m0 + row, gc = n0 + col;
        const double full = acc64[e];
        T val = static_cast<T>(full);
        if (EPI == kEpiSpectral) {
            // P = F / (4 nx ny) / (u_j^2 + v_k^2), P[0][0] = 0 (cpu_poisson.cpp:163-173)
            if (gr < epi.nx && gc < epi.ny) {
                const T uj = epi.u[gr], vk = epi.v[gc];
                const T den = uj * uj + vk * vk;
                val = den > T(0) ? static_cast<T>(full * static_cast<double>(epi.invNorm) /
                                                  static_cast<double>(den))
                                 : T(0);
            } else {
                val = T(0);
            }
        }
        if (TRANS)
            C[(static_cast<long>(gc) * ldc + gr) * cstep] = val;
        else
            C[(static_cast<long>(gr) * ldc + gc) * cstep] = val;
    }
}

template <typename T, int EPI, bool TRANS>
void gemmDispatch(int batch, const GemmBatch<T>& gb, int M, int N, int K, int lda, int ldb,
                  int ldc, const SpectralEpi<T>& epi, hipStream_t s, const int* skip = nullptr,
                  int cstep = 1) {
    if (M % 64 || N % 64 || K % 64) throw std::runtime_error("gemm: dims must be multiples of 64");
    if (batch < 1 || batch > 4) throw std::runtime_error("gemm: batch must be 1..4");
    static bool attrSet = false;   // the panels need more than the default 64 KB of LDS
    constexpr unsigned shm = 2 * PanelGeom<T>::ROWS * PanelGeom<T>::LD * sizeof(T);
    if (!attrSet) {
        SPHIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_gemm<T, EPI, TRANS>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, shm));
        attrSet = true;
    }
    dim3 grid(N / 32, M / 32, batch);
    launchKernel(k_gemm<T, EPI, TRANS>, grid, 256, shm, s, gb, M, N, K, lda, ldb, ldc, cstep, epi,
                 skip);
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
    gemmDispatch<T, kEpiNone, false>(batch, gb, M, N, K, lda, ldb, ldc, SpectralEpi<T>{}, stream);
}

template <typename T>
PoissonPlan<T>::PoissonPlan(int nx, int ny, double stepX, double stepY, hipStream_t stream,
                            bool interleaved)
    : nx_(nx), ny_(ny), mx_(roundUp(nx, 64)), my_(roundUp(ny, 64)), stream_(stream),
      interleaved_(interleaved) {
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
    // The products take their right operand transposed (k contiguous), so the
    // y-axis matrices are stored as [output index][summed index].
    for (int y = 0; y < ny; ++y)
        for (int k = 0; k < ny; ++k) {
            // forward REDFT10 along y: T0[x][k] = sum_y rho[x][y] * 2 cos(pi k (y+1/2)/ny)
            cyT[static_cast<size_t>(k) * my_ + y] = 2.0 * std::cos(pi * k * (y + 0.5) / ny);
            // inverse REDFT01 along y: T1[j][y] = sum_k P[j][k] c_k cos(pi k (y+1/2)/ny)
            ayT[static_cast<size_t>(y) * my_ + k] =
                (k == 0 ? 1.0 : 2.0) * std::cos(pi * k * (y + 0.5) / ny);
            // RODFT01 along y of P shifted left and scaled by v (cpu_poisson.cpp:197-206):
            // T2[j][y] = sum_{k>=1} P[j][k] 2 v_k sin(pi k (y+1/2)/ny)
            bsv[static_cast<size_t>(y) * my_ + k] =
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
    for (T** b : {&rho_, &phi_, &t0_, &p_, &t1_, &t2_}) {
        *b = devAlloc<T>(gx);
        SPHIP_CHECK(hipMemsetAsync(*b, 0, gx * sizeof(T), stream));
    }
    if (interleaved) {
        exy_ = devAlloc<T>(2 * gx);
        SPHIP_CHECK(hipMemsetAsync(exy_, 0, 2 * gx * sizeof(T), stream));
    } else {
        for (T** b : {&ex_, &ey_}) {
            *b = devAlloc<T>(gx);
            SPHIP_CHECK(hipMemsetAsync(*b, 0, gx * sizeof(T), stream));
        }
    }
    SPHIP_CHECK(hipStreamSynchronize(stream));
}

template <typename T>
PoissonPlan<T>::~PoissonPlan() {
    for (T* b : {rho_, phi_, ex_, ey_, exy_, t0_, p_, t1_, t2_, cyT_, cx_, ayT_, bsv_, sxu_, ax_, u_,
                 v_})
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
    // Intermediates alternate between transposed and plain layouts so that
    // every product reads both operands along k: T0^T, P, T1^T / T2^T, then
    // the fields in the plain [x][y] layout.
    {   // T0^T = (rho * Cy^T)^T
        GemmBatch<T> gb{};
        gb.A[0] = rho_; gb.B[0] = cyT_; gb.C[0] = t0_;
        gemmDispatch<T, kEpiNone, true>(1, gb, M, N, N, N, N, M, none, stream_, skip);
    }
    {   // P = spectral(Cx * T0)
        GemmBatch<T> gb{};
        gb.A[0] = cx_; gb.B[0] = t0_; gb.C[0] = p_;
        SpectralEpi<T> epi{u_, v_, nx_, ny_, invNorm_};
        gemmDispatch<T, kEpiSpectral, false>(1, gb, M, N, M, M, M, N, epi, stream_, skip);
    }
    {   // T1^T = (P * Ay^T)^T, T2^T = (P * Bsv)^T
        GemmBatch<T> gb{};
        gb.A[0] = p_; gb.B[0] = ayT_; gb.C[0] = t1_;
        gb.A[1] = p_; gb.B[1] = bsv_; gb.C[1] = t2_;
        gemmDispatch<T, kEpiNone, true>(2, gb, M, N, N, N, N, M, none, stream_, skip);
    }
    {   // xi_x = Sxu * T1, xi_y = Ax * T2 (interleaved or not), phi = Ax * T1
        GemmBatch<T> gb{};
        gb.A[0] = sxu_; gb.B[0] = t1_; gb.C[0] = interleaved_ ? exy_ : ex_;
        gb.A[1] = ax_;  gb.B[1] = t2_; gb.C[1] = interleaved_ ? exy_ + 1 : ey_;
        gemmDispatch<T, kEpiNone, false>(2, gb, M, N, M, M, M, N, none, stream_, skip,
                                         interleaved_ ? 2 : 1);
        if (withPhi) {
            GemmBatch<T> gp{};
            gp.A[0] = ax_; gp.B[0] = t1_; gp.C[0] = phi_;
            gemmDispatch<T, kEpiNone, false>(1, gp, M, N, M, M, M, N, none, stream_, skip);
        }
    }
}

template class PoissonPlan<float>;
template class PoissonPlan<double>;
template void launchGemm<float>(int, const float* const*, const float* const*, float* const*,
                                int, int, int, int, int, int, hipStream_t);
template void launchGemm<double>(int, const double* const*, const double* const*,
                                 double* const*, int, int, int, int, int, int, hipStream_t);

}  // namespace sphip
