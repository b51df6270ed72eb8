// DCT Poisson solve as dense transforms on MFMA (CDNA4 / gfx950).
//
// Math: src/numeric/cpu/cpu_poisson.cpp:3-40 (FFTW REDFT10/REDFT01/RODFT01
// conventions, spectral divide, index shift for the sine transforms).  Here
// every axis transform is a matrix product on the matrix cores:
// v_mfma_f32_16x16x4_f32 (FP32 path) or v_mfma_f64_16x16x4_f64 (FP64 path),
// LDS-staged 64-wide wavefront tiles.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

#include "poisson.h"
#include "wa_device.h"
#include "density_device.h"

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

// ------------------------------------------------------------ split-BF16 products
// C = A * Bt^T as k_gemm<float>, on the BF16 matrix cores: every FP32
// operand x is split into hi = bf16(x) and lo = bf16(x - hi) (round to
// nearest even; |x - hi - lo| <= 2^-17 |x| roughly), and each product is
// a_hi b_hi + a_hi b_lo + a_lo b_hi: three v_mfma_f32_16x16x32_bf16 per 32 k,
// accumulated in FP32 over a panel of 256 k, then in FP64.  The dropped
// a_lo b_lo and the split residues leave about 2^-16 relative error per
// product.  Lane layout of the 16x16x32 instruction (measured on gfx950):
// lane l supplies A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16],
// j < 8, and holds C[4 (l / 16) + e][l % 16], e < 4, as the f32 16x16x4 form.
// The panels are split while they are copied into LDS (hi and lo arrays,
// 32 rows of 256 k, rows padded by 16 bytes).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

struct X3Geom {
    static constexpr int KP = 256;        // panel k columns
    static constexpr int LD = KP + 8;     // padded row, in bf16
    static constexpr int ROWS = 32;
    static constexpr unsigned kLds = 4u * ROWS * LD * sizeof(__bf16);   // A hi, A lo, B hi, B lo
};

__device__ inline void splitToLds(const float4& v, __bf16* hi, __bf16* lo) {
    bf16x4 h, l;
    const float x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        h[i] = static_cast<__bf16>(x[i]);
        l[i] = static_cast<__bf16>(x[i] - static_cast<float>(h[i]));
    }
    *reinterpret_cast<bf16x4*>(hi) = h;
    *reinterpret_cast<bf16x4*>(lo) = l;
}

template <int EPI, bool TRANS>
__global__ __launch_bounds__(256) void k_gemm_x3(GemmBatch<float> batch, int M, int N, int K,
                                                 int lda, int ldb, int ldc, int cstep,
                                                 SpectralEpi<float> epi, const int* skip) {
    if (skip && *skip) return;   // device loop finished: uniform early exit
    using G = X3Geom;
    constexpr int KP = G::KP, LD = G::LD, ROWS = G::ROWS;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __bf16* Ah = reinterpret_cast<__bf16*>(smem);
    __bf16* Al = Ah + ROWS * LD;
    __bf16* Bh = Al + ROWS * LD;
    __bf16* Bl = Bh + ROWS * LD;

    const float* __restrict__ A = batch.A[blockIdx.z];
    const float* __restrict__ Bt = batch.B[blockIdx.z];
    float* __restrict__ C = batch.C[blockIdx.z];
    const int m0 = blockIdx.y * ROWS, n0 = blockIdx.x * ROWS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int wm = wave >> 1, wn = wave & 1;
    const int aOff = (wm * 16 + r) * LD + 8 * q, bOff = (wn * 16 + r) * LD + 8 * q;

    double acc64[4] = {0.0, 0.0, 0.0, 0.0};
    // each thread copies 8 float4 of each operand per full panel
    constexpr int kPer = ROWS * KP / 4 / 256;
    for (int kc = 0; kc < K; kc += KP) {
        const int kw = K - kc < KP ? K - kc : KP;   // multiple of 64
        const int v4 = kw / 4;                      // float4 per panel row
        if (kc > 0) __syncthreads();                // previous panel fully read
        float4 av[kPer], bv[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int e = threadIdx.x + 256 * i, row = e / v4, c4 = e - row * v4;
            if (row < ROWS) {
                av[i] = *reinterpret_cast<const float4*>(A + static_cast<long>(m0 + row) * lda +
                                                         kc + 4 * c4);
                bv[i] = *reinterpret_cast<const float4*>(Bt + static_cast<long>(n0 + row) * ldb +
                                                         kc + 4 * c4);
            }
        }
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int e = threadIdx.x + 256 * i, row = e / v4, c4 = e - row * v4;
            if (row < ROWS) {
                splitToLds(av[i], Ah + row * LD + 4 * c4, Al + row * LD + 4 * c4);
                splitToLds(bv[i], Bh + row * LD + 4 * c4, Bl + row * LD + 4 * c4);
            }
        }
        __syncthreads();
        vec4<float> acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < kw; k0 += 32) {
            const bf16x8 ah = *reinterpret_cast<const bf16x8*>(Ah + aOff + k0);
            const bf16x8 al = *reinterpret_cast<const bf16x8*>(Al + aOff + k0);
            const bf16x8 bh = *reinterpret_cast<const bf16x8*>(Bh + bOff + k0);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(Bl + bOff + k0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) acc64[e] += static_cast<double>(acc[e]);
    }

#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int row = wm * 16 + 4 * q + e;
        const int col = wn * 16 + r;
        const int gr = m0 + row, gc = n0 + col;
        const double full = acc64[e];
        float val = static_cast<float>(full);
        if (EPI == kEpiSpectral) {
            // P = F / (4 nx ny) / (u_j^2 + v_k^2), P[0][0] = 0 (cpu_poisson.cpp:163-173)
            if (gr < epi.nx && gc < epi.ny) {
                const float uj = epi.u[gr], vk = epi.v[gc];
                const float den = uj * uj + vk * vk;
                val = den > 0.f ? static_cast<float>(full * static_cast<double>(epi.invNorm) /
                                                     static_cast<double>(den))
                                : 0.f;
            } else {
                val = 0.f;
            }
        }
        if (TRANS)
            C[(static_cast<long>(gc) * ldc + gr) * cstep] = val;
        else
            C[(static_cast<long>(gr) * ldc + gc) * cstep] = val;
    }
}

// ------------------------------------------------------------ FFT path
// For power-of-two grids the three axis passes run as complex FFTs in LDS
// (Makhoul's mapping of the cosine/sine transforms onto a length-L complex
// FFT; J. Makhoul, "A fast cosine transform in one and two dimensions",
// IEEE TASSP 28(1), 1980), so a solve moves each grid a few times through
// HBM instead of running dense N^3 products:
//   pass A (rows, along y):    F1 = REDFT10_y(rho)
//   pass B (columns, along x): P  = spectral(REDFT10_x(F1)),
//                              Hc = REDFT01_x(P), Hs = RODFT01_x(u P shifted)
//   pass C (rows, along y):    xi_x = REDFT01_y(Hs), xi_y = RODFT01_y(v Hc shifted),
//                              phi = REDFT01_y(Hc)
// Two real lines share one complex transform (one as the real part, one as
// the imaginary part).  For a real line x of length L:
//   REDFT10:  z[n] = x[2n], z[L-1-n] = x[2n+1]; Z = FFT(z);
//             X[k] = 2 Re(exp(-i pi k / 2L) Z[k])
//   REDFT01 (y[n] = a[0] + 2 sum_k a[k] cos(pi k (n+1/2)/L)):
//             Z[k] = (a[k] - i a[L-k]) exp(i pi k / 2L) (a[L] = 0) is
//             Hermitian, so z = IFFT(Z) is real; y[2m] = z[m], y[2m+1] = z[L-1-m]
//   the shifted sine transform y[n] = sum_{k>=1} 2 b[k] sin(pi k (n+1/2)/L)
//             equals (-1)^n REDFT01(a)[n] with a[m] = b[L-m], a[0] = 0.
// One team of threads owns the transforms of its two lines (one row in the
// last pass);
// the FFT is a Stockham autosort (radix 8 or 4, a final radix 2 where
// needed) between two LDS buffers, with twiddles exp(-2 pi i t / L) from an
// LDS table.

template <typename T>
using cplx = typename Vec2<T>::type;

template <typename T>
__device__ inline cplx<T> cmk(T re, T im) {
    cplx<T> c;
    c.x = re;
    c.y = im;
    return c;
}
template <typename T>
__device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return cmk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return cmk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) { return cmk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ inline cplx<T> cconj(cplx<T> a) { return cmk<T>(a.x, -a.y); }

// Position of element i of a line in the REDFT10 reordering (and of output
// element n of REDFT01 in z).
__device__ inline int dctPos(int i, int L) { return (i & 1) ? L - 1 - (i >> 1) : (i >> 1); }

// Radix-4 DFT in registers: forward uses exp(-2 pi i/4) = -i, inverse +i.
template <typename T, bool INV>
__device__ inline void dft4(cplx<T>& p0, cplx<T>& p1, cplx<T>& p2, cplx<T>& p3) {
    const cplx<T> t0 = cadd<T>(p0, p2), t1 = csub<T>(p0, p2), t2 = cadd<T>(p1, p3);
    const cplx<T> d = csub<T>(p1, p3);
    const cplx<T> t3 = INV ? cmk<T>(-d.y, d.x) : cmk<T>(d.y, -d.x);
    p0 = cadd<T>(t0, t2);
    p1 = cadd<T>(t1, t3);
    p2 = csub<T>(t0, t2);
    p3 = csub<T>(t1, t3);
}

// Radix-8 DFT in registers: a radix-2 step between v[r] and v[r+4] (the
// differences turned by exp(-+2 pi i r/8)), then two radix-4 DFTs giving the
// even and the odd outputs.
template <typename T, bool INV>
__device__ inline void dft8(cplx<T>* v) {
    constexpr T h = T(0.70710678118654752440);
    cplx<T> u0 = cadd<T>(v[0], v[4]), u1 = cadd<T>(v[1], v[5]), u2 = cadd<T>(v[2], v[6]),
            u3 = cadd<T>(v[3], v[7]);
    cplx<T> d0 = csub<T>(v[0], v[4]), d1 = csub<T>(v[1], v[5]), d2 = csub<T>(v[2], v[6]),
            d3 = csub<T>(v[3], v[7]);
    // forward: d1 (1 - i)/sqrt2, d2 (-i), d3 (-1 - i)/sqrt2; inverse: conjugates
    if (INV) {
        d1 = cmk<T>(h * (d1.x - d1.y), h * (d1.x + d1.y));
        d2 = cmk<T>(-d2.y, d2.x);
        d3 = cmk<T>(-h * (d3.x + d3.y), h * (d3.x - d3.y));
    } else {
        d1 = cmk<T>(h * (d1.x + d1.y), h * (d1.y - d1.x));
        d2 = cmk<T>(d2.y, -d2.x);
        d3 = cmk<T>(h * (d3.y - d3.x), -h * (d3.x + d3.y));
    }
    dft4<T, INV>(u0, u1, u2, u3);
    dft4<T, INV>(d0, d1, d2, d3);
    v[0] = u0; v[1] = d0; v[2] = u1; v[3] = d1;
    v[4] = u2; v[5] = d2; v[6] = u3; v[7] = d3;
}

// exp(-2 pi i t / 4L), t < 4L, as the product of two small tables:
// lo[t mod 2^S] hi[t >> S] with S = ceil(log2(4L) / 2), together at most 128
// entries for L <= 1024 (a full table of L twiddles and L half-angle
// factors read by every block of a pass made its loads the slowest phase
// of the pass).  The FFT twiddle exp(-2 pi i t / L) is entry 4t, the
// half-angle factor exp(-i pi k / 2L) of the cosine transforms entry k.
// The tables hold FP64 values and the product is taken in FP64, so an FP32
// factor is the FP64 value rounded once, as a full FP32 table holds it.
__host__ __device__ constexpr int twShift(int L) {
    int m = 0;
    while ((1 << m) < 4 * L) ++m;
    return (m + 1) / 2;
}
__host__ __device__ constexpr int twLo(int L) { return 1 << twShift(L); }
__host__ __device__ constexpr int twCount(int L) { return twLo(L) + ((4 * L) >> twShift(L)); }

template <typename T>
struct Twid {
    const cplx<double>* fac;   // lo (nLo entries), then hi
    int s, nLo;
    __device__ cplx<T> at(int t) const {
        const cplx<double> w = cmul<double>(fac[nLo + (t >> s)], fac[t & (nLo - 1)]);
        return cmk<T>(static_cast<T>(w.x), static_cast<T>(w.y));
    }
    __device__ cplx<T> tw(int t) const { return at(4 * t); }
    __device__ cplx<T> hf(int k) const { return at(k); }
};

// One Stockham stage of radix R for K transforms at once, butterflies spread
// over the block's threads: butterfly j reads a[j + r L/R], turns input r by
// tw[r (j mod ns) L/(ns R)] and writes b[(j - j mod ns) R + j mod ns + r ns].
template <typename T, bool INV, int R, int K>
__device__ inline void fftStage(cplx<T>* const* a, cplx<T>* const* b, const Twid<T>& tw, int L,
                                int ns, int tid, int nt) {
    const int nb = L / R, st = L / (ns * R);
    for (int j = tid; j < nb; j += nt) {
        const int jm = j & (ns - 1);
        cplx<T> v[K][R];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int r = 0; r < R; ++r) v[k][r] = a[k][j + r * nb];
        if (jm) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cplx<T> w = tw.tw(r * jm * st);
                if (INV) w = cconj<T>(w);
#pragma unroll
                for (int k = 0; k < K; ++k) v[k][r] = cmul<T>(v[k][r], w);
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if constexpr (R == 8) {
                dft8<T, INV>(v[k]);
            } else if constexpr (R == 4) {
                dft4<T, INV>(v[k][0], v[k][1], v[k][2], v[k][3]);
            } else {
                const cplx<T> t = v[k][0];
                v[k][0] = cadd<T>(t, v[k][1]);
                v[k][1] = csub<T>(t, v[k][1]);
            }
        }
        const int d = (j - jm) * R + jm;
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int r = 0; r < R; ++r) b[k][d + r * ns] = v[k][r];
    }
}

// Stockham FFT of length L (power of two, >= 2) by a team of nt threads
// (tid = the thread's index in its team), K transforms of the same length at
// once: a[k] is transform k's input in natural order, b[k] its scratch; on
// return a[k] points at its result.  Each stage takes the largest radix
// (8, 4) that divides what is left and still gives every thread a
// butterfly, else the largest that divides (a 512-point transform on 128
// threads: radix 4, 4, 4, 4, 2).  INV: the unnormalised inverse
// (exp(+2 pi i nk/L)).  Stages end in block barriers, so every team of the
// block runs the same L.
template <typename T, bool INV, int K>
__device__ void blockFft(cplx<T>** a, cplx<T>** b, const Twid<T>& tw, int L, int tid, int nt) {
    for (int ns = 1; ns < L;) {
        const int q = L / ns;
        int R;
        if ((q & 7) == 0 && L / 8 >= nt) R = 8;
        else if ((q & 3) == 0 && L / 4 >= nt) R = 4;
        else R = (q & 7) == 0 ? 8 : (q & 3) == 0 ? 4 : 2;
        if (R == 8) fftStage<T, INV, 8, K>(a, b, tw, L, ns, tid, nt);
        else if (R == 4) fftStage<T, INV, 4, K>(a, b, tw, L, ns, tid, nt);
        else fftStage<T, INV, 2, K>(a, b, tw, L, ns, tid, nt);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            cplx<T>* t = a[k];
            a[k] = b[k];
            b[k] = t;
        }
        ns *= R;
    }
}

// LDS layout of the FFT passes: the twiddle factor tables (Twid), then the
// waves' buffers.
template <typename T>
struct FftLds {
    cplx<double>* fac;
    cplx<T>* waves;
    __device__ Twid<T> twid(int L) const { return Twid<T>{fac, twShift(L), twLo(L)}; }
};
// Complex T entries of LDS ahead of the waves' buffers: the FP64 factor tables.
template <typename T>
constexpr int kFftLdsHead(int L) {
    return twCount(L) * static_cast<int>(sizeof(cplx<double>) / sizeof(cplx<T>));
}
template <typename T>
__device__ inline FftLds<T> fftLds(unsigned char* smem, int L) {
    FftLds<T> f;
    f.fac = reinterpret_cast<cplx<double>*>(smem);
    f.waves = reinterpret_cast<cplx<T>*>(smem) + kFftLdsHead<T>(L);
    return f;
}

// Copies the factor tables (at most 128 entries) into LDS; the caller
// synchronises before use.
template <typename T>
__device__ inline void loadTables(const FftLds<T>& f, const cplx<double>* gfac, int L) {
    const int n = twCount(L);
    for (int t = threadIdx.x; t < n; t += blockDim.x) f.fac[t] = gfac[t];
}

// REDFT10 outputs for the pair (k, L-k) from the transform Z of two packed
// lines: X_a, X_b at k and (when 0 < k < L/2) at L-k.
template <typename T>
__device__ inline void dct2Pair(const cplx<T>* Z, const Twid<T>& tw, int k, int L, T* xa, T* xb,
                                T* xaM, T* xbM) {
    const cplx<T> zk = Z[k], zm = cconj<T>(Z[(L - k) & (L - 1)]);
    const cplx<T> va = cmk<T>(T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y + zm.y));
    const cplx<T> dv = csub<T>(zk, zm);
    const cplx<T> vb = cmk<T>(T(0.5) * dv.y, T(-0.5) * dv.x);   // -i (Z[k] - conj Z[L-k]) / 2
    const cplx<T> h = tw.hf(k);
    *xa = T(2) * (h.x * va.x - h.y * va.y);
    *xb = T(2) * (h.x * vb.x - h.y * vb.y);
    // at L-k: V[L-k] = conj(V[k]) and exp(-i pi (L-k)/2L) = -i conj(h)
    const cplx<T> hm = cmk<T>(-h.y, -h.x);
    *xaM = T(2) * (hm.x * va.x + hm.y * va.y);
    *xbM = T(2) * (hm.x * vb.x + hm.y * vb.y);
}

// REDFT01 input pair: Z[j] and Z[L-j] of the line a (a[L] = 0) given a[j] and
// a[L-j] (for j = 0 pass aM = 0).
template <typename T>
__device__ inline void dct3InPair(T aj, T aM, cplx<T> h, cplx<T>* zj, cplx<T>* zM) {
    // exp(+i pi j / 2L) = conj(h); exp(+i pi (L-j)/2L) = i h
    const cplx<T> g = cconj<T>(h), gm = cmk<T>(-h.y, h.x);
    *zj = cmul<T>(cmk<T>(aj, -aM), g);
    *zM = cmul<T>(cmk<T>(aM, -aj), gm);
}

// An axis: its FP64 factor tables in device memory (lo, then hi; see
// Twid), its length and bin step.
template <typename T>
struct FftTab {
    const cplx<double>* fac;
    int n;
    double step;
};

// The spectral scale of frequency j along an axis, u_j = pi j / (n step)
// (cpu_poisson.cpp:163-173), computed where it is used, in the arithmetic
// of the plan's scale tables: the reference's float arithmetic in FP32,
// double in FP64.
template <typename T>
__device__ inline T specScale(int j, const FftTab<T>& a);
template <>
__device__ inline float specScale<float>(int j, const FftTab<float>& a) {
    return static_cast<float>(3.14159265358979323846f) * static_cast<float>(j) /
           (static_cast<float>(a.n) * static_cast<float>(a.step));
}
template <>
__device__ inline double specScale<double>(int j, const FftTab<double>& a) {
    return 3.14159265358979323846 * j / (a.n * a.step);
}

// Global loads are issued kFftBatch at a time into registers before their
// LDS stores, so their latencies overlap.
constexpr int kFftBatch = 8;

// Threads per block of the FFT passes: one block transforms one line pair
// (passes A, B) or one row (pass C); two waves for lines of 512 and more.
__host__ __device__ inline int fftThreads(int L) { return L >= 512 ? 128 : 64; }

// Pass A: REDFT10 along y of row pairs, one pair per team of
// nt = fftThreads(ny) threads: team q of block `block` takes rows
// 2 (block * teams + q) and the next.
template <typename T>
__device__ inline void dctRowsBlock(const T* __restrict__ rho, T* __restrict__ out, int nx,
                                    int ny, int ld, const FftTab<T>& tab, int block) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int L = ny, nt = fftThreads(L), teams = blockDim.x / nt;
    const int team = threadIdx.x / nt, tid = threadIdx.x - team * nt;
    const FftLds<T> f = fftLds<T>(smem, L);
    cplx<T>* a[1] = {f.waves + static_cast<long>(team) * 2 * L};
    cplx<T>* b[1] = {a[0] + L};
    loadTables<T>(f, tab.fac, L);
    const Twid<T> tw = f.twid(L);
    const int r0 = 2 * (block * teams + team);
    const T* xa = rho + static_cast<long>(r0) * ld;
    const T* xb = xa + ld;
    for (int i0 = tid; i0 < L; i0 += nt * kFftBatch) {
        T va[kFftBatch], vb[kFftBatch];
#pragma unroll
        for (int q = 0; q < kFftBatch; ++q) {
            const int i = i0 + nt * q;
            va[q] = i < L ? xa[i] : T(0);
            vb[q] = i < L ? xb[i] : T(0);
        }
#pragma unroll
        for (int q = 0; q < kFftBatch; ++q) {
            const int i = i0 + nt * q;
            if (i < L) a[0][dctPos(i, L)] = cmk<T>(va[q], vb[q]);
        }
    }
    __syncthreads();
    blockFft<T, false, 1>(a, b, tw, L, tid, nt);
    const cplx<T>* Z = a[0];
    T* oa = out + static_cast<long>(r0) * ld;
    T* ob = oa + ld;
    for (int k = tid; k <= L / 2; k += nt) {
        T ra, rb, raM, rbM;
        dct2Pair<T>(Z, tw, k, L, &ra, &rb, &raM, &rbM);
        oa[k] = ra;
        ob[k] = rb;
        if (k > 0 && k < L / 2) {
            oa[L - k] = raM;
            ob[L - k] = rbM;
        }
    }
    (void)nx;
}

// Pass A reading the integer density grids instead of rho, one row pair
// per block of kRhoThreads threads: the block's threads finalize the pair's
// bins (finalizeBin: densities, rho, tau term, clearing) straight into the
// transform input, kRhoThreads consecutive bins per round, then the first
// fftThreads(ny) threads run the transform.  The launch clears the tile
// counters.
constexpr int kRhoThreads = 2 * kBlock;
constexpr int kRhoRounds = 4;   // rows of at most 1024 bins

// The tau partials of the fused pass.  Round r's bins are 2 kBlock
// consecutive bins; each half of the block reduces its kBlock bins with the
// finalize kernel's tree (the same wave shuffles, waves summed in the same
// order), so partial first + 2 r + half holds the bits the finalize kernel
// writes for the same bins.
__device__ inline void halfBlockPartials(double (&v)[kRhoRounds], int rounds, double* partials,
                                         long first) {
    constexpr int W = kBlock / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int j = 0; j < kRhoRounds; ++j) v[j] += __shfl_down(v[j], off, 64);
    __shared__ double sh[kRhoThreads / 64][kRhoRounds];
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < kRhoRounds; ++j) sh[wave][j] = v[j];
    __syncthreads();
    const int half = threadIdx.x / kBlock;
    if (threadIdx.x % kBlock == 0)
        for (int j = 0; j < rounds; ++j) {
            double s = 0;
            for (int w = 0; w < W; ++w) s += sh[half * W + w][j];
            partials[first + 2 * j + half] = s;
        }
}

template <typename T>
__global__ __launch_bounds__(kRhoThreads) void k_dct_rows_rho(DensityIn<T> in, T* __restrict__ out,
                                                              int nx, int ny, int ld,
                                                              FftTab<T> tab, const int* skip) {
    if (skip && *skip) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int L = ny, nt = fftThreads(L), rounds = 2 * L / kRhoThreads;
    const FftLds<T> f = fftLds<T>(smem, L);
    cplx<T>* a[1] = {f.waves};
    cplx<T>* b[1] = {f.waves + L};
    loadTables<T>(f, tab.fac, L);
    const Twid<T> tw = f.twid(L);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < in.tg.nKeys; t += gridDim.x * blockDim.x)
        clearTileKeys(in.tg, t);
    const int r0 = 2 * blockIdx.x;
    double over[kRhoRounds] = {0.0, 0.0, 0.0, 0.0};
    // rounds in pairs: both rounds' loads in flight before the first
    // clearing store
#pragma unroll
    for (int j0 = 0; j0 < kRhoRounds; j0 += 2) {
        if (j0 >= rounds) break;
        BinLoads bl[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (j0 + h < rounds)
                finalizeLoad<T>(in, r0 * L + (j0 + h) * kRhoThreads + threadIdx.x, bl[h]);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int j = j0 + h;
            if (j < rounds) {
                const int e = j * kRhoThreads + threadIdx.x;
                const int row = e / L, i = e - row * L;
                const T val = finalizeFinish<T>(in, r0 * L + e, bl[h], &over[j]);
                reinterpret_cast<T*>(a[0] + dctPos(i, L))[row] = val;
            }
        }
    }
    __syncthreads();
    // threads past the team take no butterflies (their index is past every
    // stage's count) but meet every barrier
    const int tid = static_cast<int>(threadIdx.x) < nt ? static_cast<int>(threadIdx.x) : L;
    blockFft<T, false, 1>(a, b, tw, L, tid, nt);
    const cplx<T>* Z = a[0];
    T* oa = out + static_cast<long>(r0) * ld;
    T* ob = oa + ld;
    for (int k = tid; k <= L / 2; k += nt) {
        T ra, rb, raM, rbM;
        dct2Pair<T>(Z, tw, k, L, &ra, &rb, &raM, &rbM);
        oa[k] = ra;
        ob[k] = rb;
        if (k > 0 && k < L / 2) {
            oa[L - k] = raM;
            ob[L - k] = rbM;
        }
    }
    halfBlockPartials(over, rounds, in.partials, static_cast<long>(blockIdx.x) * 2 * rounds);
    (void)nx;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dct_rows(const T* __restrict__ rho, T* __restrict__ out,
                                                     int nx, int ny, int ld, FftTab<T> tab,
                                                     const int* skip) {
    if (skip && *skip) return;
    dctRowsBlock<T>(rho, out, nx, ny, ld, tab, blockIdx.x);
}

// Pass B: G column pairs per block (pair g: columns c0 + 2g, c0 + 2g + 1),
// one team of nt = blockDim.x / G threads each: REDFT10 along x, the
// spectral divide (cpu_poisson.cpp:163-173), then both inverse transforms
// along x at once (cosine -> Hc, shifted sine of u P -> Hs).  The block
// loads and stores 2G adjacent values per row.
// Three line buffers per team: the two inverse transforms
// run one after the other, sharing the third buffer (two at once need a
// fourth buffer and, holding both transforms' butterflies, spill registers
// in a 256-thread block), which also leaves LDS room for the WA blocks that
// share the fused launch.
template <typename T>
__device__ inline void dctColsBlock(const T* __restrict__ F1, T* __restrict__ Hc,
                                    T* __restrict__ Hs, int G, int nx, int ny, int ld,
                                    const FftTab<T>& tab, const FftTab<T>& taby, T invNorm,
                                    int block) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int L = nx, nt = blockDim.x / G;
    const int team = threadIdx.x / nt, tid = threadIdx.x - team * nt;
    const FftLds<T> f = fftLds<T>(smem, L);
    loadTables<T>(f, tab.fac, L);
    const Twid<T> tw = f.twid(L);
    // team g owns region[NB L g, NB L (g + 1))
    constexpr int NB = 3;
    cplx<T>* region = f.waves;
    cplx<T>* b0 = region + static_cast<long>(team) * NB * L;
    cplx<T>* b1 = b0 + L;
    cplx<T>* b2 = b1 + L;
    const int C = 2 * G, c0 = block * C;
    // element e: row e / C, column c0 + e % C (C adjacent values per row)
    const int total = L * C;
    for (int e0 = threadIdx.x; e0 < total; e0 += blockDim.x * kFftBatch) {
        T vals[kFftBatch];
#pragma unroll
        for (int q = 0; q < kFftBatch; ++q) {
            const int e = e0 + blockDim.x * q;
            const int x = e / C, c = e - x * C;
            vals[q] = e < total ? F1[static_cast<long>(x) * ld + c0 + c] : T(0);
        }
#pragma unroll
        for (int q = 0; q < kFftBatch; ++q) {
            const int e = e0 + blockDim.x * q;
            if (e < total) {
                const int x = e / C, c = e - x * C;
                reinterpret_cast<T*>(region + static_cast<long>(c >> 1) * NB * L + dctPos(x, L))[c & 1] =
                    vals[q];
            }
        }
    }
    const int ka = c0 + 2 * team;
    const T va = specScale<T>(ka, taby), vb = specScale<T>(ka + 1, taby);
    __syncthreads();
    cplx<T>* fa[1] = {b0};
    cplx<T>* fb[1] = {b1};
    blockFft<T, false, 1>(fa, fb, tw, L, tid, nt);
    cplx<T>* r = fa[0];
    cplx<T>* o = fb[0];
    // each thread owns the pairs (j, L-j): REDFT10 outputs, spectral divide,
    // then the inputs of the two inverse transforms (W2 in place in r, W1 in o)
    for (int j = tid; j <= L / 2; j += nt) {
        const int jm = (L - j) & (L - 1);
        T ga, gb, gaM, gbM;
        dct2Pair<T>(r, tw, j, L, &ga, &gb, &gaM, &gbM);
        if (j == 0 || j == L / 2) {
            gaM = ga;
            gbM = gb;
        }
        const T uj = specScale<T>(j, tab), um = specScale<T>(jm, tab);
        auto spec = [&](T g, T uu, T vv) {
            const T den = uu * uu + vv * vv;
            return den > T(0) ? g * invNorm / den : T(0);
        };
        const T pa = spec(ga, uj, va), pb = spec(gb, uj, vb);
        const T paM = j == 0 ? T(0) : spec(gaM, um, va), pbM = j == 0 ? T(0) : spec(gbM, um, vb);
        const cplx<T> h = tw.hf(j);
        cplx<T> za, zaM, zb, zbM;
        dct3InPair<T>(pa, paM, h, &za, &zaM);
        dct3InPair<T>(pb, pbM, h, &zb, &zbM);
        const cplx<T> w1 = cmk<T>(za.x - zb.y, za.y + zb.x), w1M = cmk<T>(zaM.x - zbM.y, zaM.y + zbM.x);
        // shifted sine: a2[m] = u[L-m] P[L-m] (m >= 1), a2[0] = 0
        const T sa = j == 0 ? T(0) : um * paM, sb = j == 0 ? T(0) : um * pbM;
        const T saM = j == 0 ? T(0) : uj * pa, sbM = j == 0 ? T(0) : uj * pb;
        dct3InPair<T>(sa, saM, h, &za, &zaM);
        dct3InPair<T>(sb, sbM, h, &zb, &zbM);
        const cplx<T> w2 = cmk<T>(za.x - zb.y, za.y + zb.x), w2M = cmk<T>(zaM.x - zbM.y, zaM.y + zbM.x);
        o[j] = w1;
        r[j] = w2;
        if (j > 0 && j < L / 2) {
            o[L - j] = w1M;
            r[L - j] = w2M;
        }
    }
    __syncthreads();
    const cplx<T>* res1;
    const cplx<T>* res2;
    cplx<T>* i1[1] = {o};
    cplx<T>* s1[1] = {b2};
    blockFft<T, true, 1>(i1, s1, tw, L, tid, nt);
    cplx<T>* i2[1] = {r};
    cplx<T>* s2[1] = {i1[0] == o ? b2 : o};
    blockFft<T, true, 1>(i2, s2, tw, L, tid, nt);
    res1 = i1[0];
    res2 = i2[0];
    // the results sit at the same offsets in every team's region
    const long off1 = res1 - b0, off2 = res2 - b0;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int x = e / C, c = e - x * C;
        const cplx<T>* w = region + static_cast<long>(c >> 1) * NB * L;
        const int pos = dctPos(x, L);
        const cplx<T> z1 = w[off1 + pos], z2 = w[off2 + pos];
        const long at = static_cast<long>(x) * ld + c0 + c;
        Hc[at] = (c & 1) ? z1.y : z1.x;
        const T sv = (c & 1) ? z2.y : z2.x;
        Hs[at] = (x & 1) ? -sv : sv;
    }
}

// Pass B's blocks own adjacent column groups whose row segments share cache
// lines; xcdBlock (common.h) gives each XCD n / 8 adjacent groups, so the
// segments of one line meet in one L2 instead of eight.

// MAXT: the largest block the instance is launched with (G teams of
// fftThreads(nx) threads); blocks of at most 256 threads take the 256-thread
// instance, whose register budget holds the transforms without spilling.
template <typename T, int MAXT>
__global__ __launch_bounds__(MAXT) void k_dct_cols(const T* __restrict__ F1, T* __restrict__ Hc,
                                                   T* __restrict__ Hs, int G, int nx, int ny,
                                                   int ld, FftTab<T> tab, FftTab<T> taby,
                                                   T invNorm, const int* skip) {
    if (skip && *skip) return;
    dctColsBlock<T>(F1, Hc, Hs, G, nx, ny, ld, tab, taby, invNorm,
                          xcdBlock(blockIdx.x, gridDim.x));
}

// Pass B with the WA pass's blocks in the same launch: blocks below
// colBlocks run the column pass (kBlock threads = G teams, three line
// buffers each so that its LDS leaves room for the WA blocks), the rest WA
// blocks waFirst, waFirst + 1, ... of a WA pass of waTotal blocks.  The WA
// pass depends only on the positions, so it fills the CUs that the column
// pass leaves idle.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dct_cols_wa(const T* __restrict__ F1,
                                                        T* __restrict__ Hc, T* __restrict__ Hs,
                                                        int G, int nx, int ny, int ld,
                                                        FftTab<T> tab, FftTab<T> taby,
                                                        T invNorm,
                                                        const int* skip, WaTask<T> wa,
                                                        int colBlocks, int waFirst, int waTotal) {
    if (skip && *skip) return;
    if (static_cast<int>(blockIdx.x) < colBlocks)
        dctColsBlock<T>(F1, Hc, Hs, G, nx, ny, ld, tab, taby, invNorm,
                               xcdBlock(blockIdx.x, colBlocks));
    else
        waBlockTask<T>(wa, waFirst + (blockIdx.x - colBlocks), waTotal);
}

// Pass C: one row per team of nt = fftThreads(ny) threads (team q of block
// `block` takes row x = block * teams + q): xi_x = REDFT01_y(Hs[x]) and
// xi_y = shifted sine of v Hc[x], packed in one complex transform, and
// phi = REDFT01_y(Hc[x]) when requested (both transforms at once).  The
// field goes to fx[e * cstep], fy[e * cstep] (cstep 2 interleaves).
template <typename T>
__device__ inline void idctRowsBlock(const T* __restrict__ Hc, const T* __restrict__ Hs,
                                     T* __restrict__ fx, T* __restrict__ fy, T* __restrict__ phi,
                                     int cstep, int nx, int ny, int ld, const FftTab<T>& tab,
                                     int block) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int L = ny, nt = fftThreads(L), teams = blockDim.x / nt;
    const int team = threadIdx.x / nt, tid = threadIdx.x - team * nt;
    const FftLds<T> f = fftLds<T>(smem, L);
    cplx<T>* a = f.waves + static_cast<long>(team) * (phi ? 4 : 2) * L;
    cplx<T>* b = a + L;
    cplx<T>* c = phi ? b + L : nullptr;
    cplx<T>* d = phi ? c + L : nullptr;
    const int x = block * teams + team;
    // the row pair (Hs[x][j], Hc[x][j]) into b; the tables load alongside
    const T* hs = Hs + static_cast<long>(x) * ld;
    const T* hc = Hc + static_cast<long>(x) * ld;
    loadTables<T>(f, tab.fac, L);
    const Twid<T> tw = f.twid(L);
    for (int i0 = tid; i0 < L; i0 += nt * kFftBatch) {
        T vsv[kFftBatch], vcv[kFftBatch];
#pragma unroll
        for (int q = 0; q < kFftBatch; ++q) {
            const int i = i0 + nt * q;
            vsv[q] = i < L ? hs[i] : T(0);
            vcv[q] = i < L ? hc[i] : T(0);
        }
#pragma unroll
        for (int q = 0; q < kFftBatch; ++q) {
            const int i = i0 + nt * q;
            if (i < L) b[i] = cmk<T>(vsv[q], vcv[q]);
        }
    }
    __syncthreads();
    for (int j = tid; j <= L / 2; j += nt) {
        const int jm = (L - j) & (L - 1);
        const cplx<T> pj = b[j], pm = j == 0 ? cmk<T>(0, 0) : b[jm];
        const cplx<T> h = tw.hf(j);
        cplx<T> za, zaM, zb, zbM;
        dct3InPair<T>(pj.x, pm.x, h, &za, &zaM);
        // shifted sine of v Hc: a2[m] = v[L-m] Hc[L-m] (m >= 1), a2[0] = 0
        const T bj = j == 0 ? T(0) : specScale<T>(jm, tab) * pm.y;
        const T bm = j == 0 ? T(0) : specScale<T>(j, tab) * pj.y;
        dct3InPair<T>(bj, bm, h, &zb, &zbM);
        a[j] = cmk<T>(za.x - zb.y, za.y + zb.x);
        if (j > 0 && j < L / 2) a[L - j] = cmk<T>(zaM.x - zbM.y, zaM.y + zbM.x);
        if (phi) {
            dct3InPair<T>(pj.y, pm.y, h, &za, &zaM);
            c[j] = za;
            if (j > 0 && j < L / 2) c[L - j] = zaM;
        }
    }
    __syncthreads();
    const long row = static_cast<long>(x) * ld;
    if (phi) {
        cplx<T>* in[2] = {a, c};
        cplx<T>* sc2[2] = {b, d};
        blockFft<T, true, 2>(in, sc2, tw, L, tid, nt);
        for (int y = tid; y < L; y += nt) phi[row + y] = in[1][dctPos(y, L)].x;
        a = in[0];
    } else {
        cplx<T>* in[1] = {a};
        cplx<T>* sc1[1] = {b};
        blockFft<T, true, 1>(in, sc1, tw, L, tid, nt);
        a = in[0];
    }
    for (int y = tid; y < L; y += nt) {
        const cplx<T> w = a[dctPos(y, L)];
        fx[(row + y) * cstep] = w.x;
        fy[(row + y) * cstep] = (y & 1) ? -w.y : w.y;
    }
    (void)nx;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_idct_rows(const T* __restrict__ Hc,
                                                      const T* __restrict__ Hs,
                                                      T* __restrict__ fx, T* __restrict__ fy,
                                                      T* __restrict__ phi, int cstep, int nx,
                                                      int ny, int ld, FftTab<T> tab,
                                                      const int* skip) {
    if (skip && *skip) return;
    idctRowsBlock<T>(Hc, Hs, fx, fy, phi, cstep, nx, ny, ld, tab, blockIdx.x);
}

template <typename T, int EPI, bool TRANS>
void gemmDispatch(int batch, const GemmBatch<T>& gb, int M, int N, int K, int lda, int ldb,
                  int ldc, const SpectralEpi<T>& epi, hipStream_t s, const int* skip = nullptr,
                  int cstep = 1, bool x3 = false) {
    if (M % 64 || N % 64 || K % 64) throw std::runtime_error("gemm: dims must be multiples of 64");
    if (batch < 1 || batch > 4) throw std::runtime_error("gemm: batch must be 1..4");
    if constexpr (std::is_same<T, float>::value) {
        if (x3) {
            constexpr unsigned shm = X3Geom::kLds;
            static std::once_flag attrOnce;
            std::call_once(attrOnce, [] {
                SPHIP_CHECK(hipFuncSetAttribute(
                    reinterpret_cast<const void*>(k_gemm_x3<EPI, TRANS>),
                    hipFuncAttributeMaxDynamicSharedMemorySize, shm));
            });
            dim3 grid(N / 32, M / 32, batch);
            launchKernel(k_gemm_x3<EPI, TRANS>, grid, 256, shm, s, gb, M, N, K, lda, ldb, ldc,
                         cstep, epi, skip);
            SPHIP_CHECK(hipGetLastError());
            return;
        }
    } else {
        if (x3) throw std::runtime_error("gemm: split-BF16 products are FP32 only");
    }
    // the panels need more than the default 64 KB of LDS; set once per
    // instantiation (several threads may solve at once)
    constexpr unsigned shm = 2 * PanelGeom<T>::ROWS * PanelGeom<T>::LD * sizeof(T);
    static std::once_flag attrOnce;
    std::call_once(attrOnce, [] {
        SPHIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_gemm<T, EPI, TRANS>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, shm));
    });
    dim3 grid(N / 32, M / 32, batch);
    launchKernel(k_gemm<T, EPI, TRANS>, grid, 256, shm, s, gb, M, N, K, lda, ldb, ldc, cstep, epi,
                 skip);
    SPHIP_CHECK(hipGetLastError());
}

template <typename T>
T* devAlloc(size_t n) {
    void* p = nullptr;
    SPHIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    g_devBytes += n * sizeof(T);
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
                int K, int lda, int ldb, int ldc, hipStream_t stream, bool x3) {
    GemmBatch<T> gb{};
    for (int i = 0; i < batch && i < 4; ++i) {
        gb.A[i] = A[i];
        gb.B[i] = B[i];
        gb.C[i] = C[i];
    }
    gemmDispatch<T, kEpiNone, false>(batch, gb, M, N, K, lda, ldb, ldc, SpectralEpi<T>{}, stream,
                                     nullptr, 1, x3);
}

template <typename T>
PoissonPlan<T>::PoissonPlan(int nx, int ny, double stepX, double stepY, hipStream_t stream,
                            bool interleaved, DevArena* arena, unsigned opts)
    : nx_(nx), ny_(ny), mx_(roundUp(nx, 64)), my_(roundUp(ny, 64)), stream_(stream),
      interleaved_(interleaved), ownsMem_(arena == nullptr), stepX_(stepX), stepY_(stepY) {
    if (nx < 2 || ny < 2) throw std::invalid_argument("Poisson grid must be at least 2x2");
    if (!(stepX > 0) || !(stepY > 0)) throw std::invalid_argument("bin step must be positive");
    if ((opts & kPlanBf16x3) && sizeof(T) != 4)
        throw std::invalid_argument("split-BF16 products are FP32 only");
    x3_ = (opts & kPlanBf16x3) != 0;
    const double pi = 3.14159265358979323846;
    const size_t gx = static_cast<size_t>(mx_) * my_;
    // an array with host contents, and a zero-filled one
    auto put = [&](T** dst, const std::vector<double>& h) {
        if (arena) arena->upload(dst, h);
        else *dst = upload<T>(h, stream);
    };
    auto zeroed = [&](T** dst, size_t count) {
        if (arena) {
            arena->zeros(dst, count);
        } else {
            *dst = devAlloc<T>(count);
            SPHIP_CHECK(hipMemsetAsync(*dst, 0, count * sizeof(T), stream));
        }
    };
    fft_ = fftShape(nx) && fftShape(ny) && !(opts & kPlanDense);
    if (fft_) {
        // the factor tables of exp(-2 pi i t / 4L) (Twid): lo, then hi
        auto tables = [&](int L) {
            const int s = twShift(L), nLo = twLo(L), n = twCount(L);
            std::vector<double> t(2 * static_cast<size_t>(n));
            for (int k = 0; k < n; ++k) {
                const double e = k < nLo ? k : static_cast<double>((k - nLo) << s);
                t[2 * k] = std::cos(2 * pi * e / (4.0 * L));
                t[2 * k + 1] = -std::sin(2 * pi * e / (4.0 * L));
            }
            return t;
        };
        if (arena) {
            arena->upload(&tabX_, tables(nx));
            arena->upload(&tabY_, tables(ny));
        } else {
            tabX_ = upload<double>(tables(nx), stream);
            tabY_ = upload<double>(tables(ny), stream);
        }
        setFftLds();
    }
    std::vector<double> u(mx_, 0.0), v(my_, 0.0);
    for (int j = 0; j < nx; ++j) u[j] = pi * j / (nx * stepX);
    for (int k = 0; k < ny; ++k) v[k] = pi * k / (ny * stepY);
    if (!fft_) {
        // Transform matrices in FP64, zero in the padding (cpu_poisson.cpp:14-36);
        // only the dense path multiplies by them.
        std::vector<double> cyT(static_cast<size_t>(my_) * my_, 0.0), ayT(cyT.size(), 0.0),
            bsv(cyT.size(), 0.0);
        std::vector<double> cx(static_cast<size_t>(mx_) * mx_, 0.0), sxu(cx.size(), 0.0),
            ax(cx.size(), 0.0);
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
        put(&cyT_, cyT);
        put(&ayT_, ayT);
        put(&bsv_, bsv);
        put(&cx_, cx);
        put(&sxu_, sxu);
        put(&ax_, ax);
    }
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
    put(&u_, uf);
    put(&v_, vf);
    invNorm_ = static_cast<T>(1.0 / (4.0 * nx * ny));
    for (T** b : {&rho_, &phi_, &t0_, &p_, &t1_, &t2_}) zeroed(b, gx);
    if (interleaved) {
        zeroed(&exy_, 2 * gx);
    } else {
        zeroed(&ex_, gx);
        zeroed(&ey_, gx);
    }
    if (!arena) SPHIP_CHECK(hipStreamSynchronize(stream));
}

template <typename T>
PoissonPlan<T>::~PoissonPlan() {
    if (!ownsMem_) return;
    for (T* b : {rho_, phi_, ex_, ey_, exy_, t0_, p_, t1_, t2_, cyT_, cx_, ayT_, bsv_, sxu_, ax_, u_,
                 v_})
        if (b) (void)hipFree(b);
    for (double* b : {tabX_, tabY_})
        if (b) (void)hipFree(b);
}

template <typename T>
double PoissonPlan<T>::flops(bool withPhi) const {
    if (fft_) return 0.0;
    const double sq_y = 2.0 * mx_ * my_ * my_, sq_x = 2.0 * mx_ * mx_ * my_;
    // forward: y then x; inverse: 2 products along y, 2 (or 3) along x
    return sq_y + sq_x + 2 * sq_y + (withPhi ? 3 : 2) * sq_x;
}

template <typename T>
double PoissonPlan<T>::bytes(bool withPhi) const {
    if (!fft_) return 0.0;
    // pass A: rho in, F1 out; pass B: F1 in, Hc and Hs out; pass C: Hc and
    // Hs in, the two field components (and phi) out
    return (9.0 + (withPhi ? 1 : 0)) * nx_ * ny_ * sizeof(T);
}

template <typename T>
bool PoissonPlan<T>::fftShape(int n) {
    return n >= 4 && n <= 1024 && (n & (n - 1)) == 0;
}

template <typename T>
void PoissonPlan<T>::setFftLds() {
    const size_t c = 2 * sizeof(T);
    ldsA_ = (kFftLdsHead<T>(ny_) + 2 * static_cast<size_t>(ny_)) * c;
    // column pairs per block of the column pass: one (so that the pass runs
    // on ny / 2 blocks; measured faster at 512^2 than gathering 4 or 8 pairs
    // per block for longer row segments), within 150 KB of LDS
    colPairs_ = 1;
    while (colPairs_ > 1 && (colPairs_ > ny_ / 2 ||
                             (kFftLdsHead<T>(nx_) + 3 * static_cast<size_t>(colPairs_) * nx_) * c >
                                 150 * 1024))
        colPairs_ >>= 1;
    ldsB_ = (kFftLdsHead<T>(nx_) + 3 * static_cast<size_t>(colPairs_) * nx_) * c;
    ldsC_ = (kFftLdsHead<T>(ny_) + 4 * static_cast<size_t>(ny_)) * c;
    // the column pass sharing its launch with the WA blocks: kBlock threads,
    // so kBlock / team column pairs per block, three line buffers per team
    // (every block of the launch reserves the same LDS, and the WA blocks
    // need the room); only where that LDS stays within kFusedLdsMax
    {
        const int nt = fftThreads(nx_), G = kBlock / nt;
        const size_t lds = (kFftLdsHead<T>(nx_) + 3 * static_cast<size_t>(G) * nx_) * c;
        ldsBW_ = (kBlock % nt == 0 && (ny_ / 2) % G == 0 && lds <= kFusedLdsMax) ? lds : 0;
        if (ldsBW_)
            raiseLdsLimit(reinterpret_cast<const void*>(k_dct_cols_wa<T>), ldsBW_);
    }
    raiseLdsLimit(reinterpret_cast<const void*>(k_dct_rows<T>), ldsA_);
    raiseLdsLimit(reinterpret_cast<const void*>(k_dct_rows_rho<T>), ldsA_);
    raiseLdsLimit(reinterpret_cast<const void*>(k_dct_cols<T, 256>), ldsB_);
    raiseLdsLimit(reinterpret_cast<const void*>(k_dct_cols<T, 1024>), ldsB_);
    raiseLdsLimit(reinterpret_cast<const void*>(k_idct_rows<T>), ldsC_);
}

template <typename T>
bool PoissonPlan<T>::setColumnPairs(int pairs) {
    if (!fft_ || pairs < 1 || (pairs & (pairs - 1)) != 0 || (ny_ / 2) % pairs != 0) return false;
    const size_t lds = (kFftLdsHead<T>(nx_) + 3 * static_cast<size_t>(pairs) * nx_) * 2 * sizeof(T);
    if (lds > 150 * 1024 || pairs * fftThreads(nx_) > 1024) return false;
    colPairs_ = pairs;
    ldsB_ = lds;
    raiseLdsLimit(reinterpret_cast<const void*>(k_dct_cols<T, 256>), ldsB_);
    raiseLdsLimit(reinterpret_cast<const void*>(k_dct_cols<T, 1024>), ldsB_);
    return true;
}

template <typename T>
void PoissonPlan<T>::solveFft(bool withPhi, const int* skip) {
    passA(nullptr, skip);
    passBC(skip, withPhi);
}

// Passes B and C of the FFT path.
template <typename T>
void PoissonPlan<T>::passBC(const int* skip, bool withPhi) {
    const FftTab<T> tx{reinterpret_cast<const cplx<double>*>(tabX_), nx_, stepX_};
    const FftTab<T> ty{reinterpret_cast<const cplx<double>*>(tabY_), ny_, stepY_};
    const int ld = my_;
    const int colThreads = colPairs_ * fftThreads(nx_);
    launchKernel(colThreads <= 256 ? k_dct_cols<T, 256> : k_dct_cols<T, 1024>,
                 dim3(ny_ / (2 * colPairs_)), dim3(colThreads), static_cast<uint32_t>(ldsB_),
                 stream_, t0_, t1_, t2_, colPairs_, nx_, ny_, ld, tx, ty, invNorm_, skip);
    SPHIP_CHECK(hipGetLastError());
    T* fx = interleaved_ ? exy_ : ex_;
    T* fy = interleaved_ ? exy_ + 1 : ey_;
    const size_t lds = withPhi ? ldsC_ : ldsC_ - 2 * static_cast<size_t>(ny_) * 2 * sizeof(T);
    launchKernel(k_idct_rows<T>, dim3(nx_), dim3(fftThreads(ny_)), static_cast<uint32_t>(lds),
                 stream_, t1_, t2_, fx, fy, withPhi ? phi_ : nullptr, interleaved_ ? 2 : 1, nx_,
                 ny_, ld, ty, skip);
    SPHIP_CHECK(hipGetLastError());
}

template <typename T>
int PoissonPlan<T>::passA(const DensityIn<T>* din, const int* skip) {
    const FftTab<T> ty{reinterpret_cast<const cplx<double>*>(tabY_), ny_, stepY_};
    if (din) {
        launchKernel(k_dct_rows_rho<T>, dim3(nx_ / 2), dim3(kRhoThreads),
                     static_cast<uint32_t>(ldsA_), stream_, *din, t0_, nx_, ny_, my_, ty, skip);
        SPHIP_CHECK(hipGetLastError());
        return (nx_ / 2) * (2 * ny_ / kBlock);
    }
    launchKernel(k_dct_rows<T>, dim3(nx_ / 2), dim3(fftThreads(ny_)),
                 static_cast<uint32_t>(ldsA_), stream_, rho_, t0_, nx_, ny_, my_, ty, skip);
    SPHIP_CHECK(hipGetLastError());
    return 0;
}

template <typename T>
bool PoissonPlan<T>::takesDensity() const {
    // the bins of a row pair split evenly into the fused pass's rounds, and
    // the transform team fits in its block
    return fft_ && nx_ == ny_ && (2 * ny_) % kRhoThreads == 0 &&
           2 * ny_ <= kRhoRounds * kRhoThreads && fftThreads(ny_) <= kRhoThreads;
}

template <typename T>
int PoissonPlan<T>::solveFromDensity(const DensityIn<T>& din, const int* skip) {
    if (!takesDensity() || din.g.dim != ny_)
        throw std::logic_error("this plan does not take the density grids");
    const int blocks = passA(&din, skip);
    passBC(skip);
    return blocks;
}

template <typename T>
int PoissonPlan<T>::fusedWaResidentBlocks() const {
    if (!fft_ || ldsBW_ == 0) return 0;
    int dev = 0, cus = 0, perCu = 0;
    SPHIP_CHECK(hipGetDevice(&dev));
    SPHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    SPHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &perCu, reinterpret_cast<const void*>(k_dct_cols_wa<T>), kBlock, ldsBW_));
    const int colBlocks = ny_ / (2 * (kBlock / fftThreads(nx_)));
    return perCu * cus - colBlocks;
}

template <typename T>
bool PoissonPlan<T>::solveWithWa(const WaTask<T>& wa, int waBlocks, const int* skip,
                                 const DensityIn<T>* din, int* densityBlocks) {
    if (!fft_ || ldsBW_ == 0 || (din && (!takesDensity() || din->g.dim != ny_))) return false;
    const FftTab<T> tx{reinterpret_cast<const cplx<double>*>(tabX_), nx_, stepX_};
    const FftTab<T> ty{reinterpret_cast<const cplx<double>*>(tabY_), ny_, stepY_};
    const int ld = my_;
    const int blocksA = passA(din, skip);
    if (densityBlocks) *densityBlocks = blocksA;
    const int G = kBlock / fftThreads(nx_), colBlocks = ny_ / (2 * G);
    launchKernel(k_dct_cols_wa<T>, dim3(colBlocks + waBlocks), dim3(kBlock),
                 static_cast<uint32_t>(ldsBW_), stream_, t0_, t1_, t2_, G, nx_, ny_, ld, tx, ty,
                 invNorm_, skip, wa, colBlocks, 0, waBlocks);
    SPHIP_CHECK(hipGetLastError());
    T* fx = interleaved_ ? exy_ : ex_;
    T* fy = interleaved_ ? exy_ + 1 : ey_;
    const size_t lds = ldsC_ - 2 * static_cast<size_t>(ny_) * 2 * sizeof(T);
    launchKernel(k_idct_rows<T>, dim3(nx_), dim3(fftThreads(ny_)), static_cast<uint32_t>(lds),
                 stream_, t1_, t2_, fx, fy, static_cast<T*>(nullptr), interleaved_ ? 2 : 1, nx_,
                 ny_, ld, ty, skip);
    SPHIP_CHECK(hipGetLastError());
    return true;
}

template <typename T>
void PoissonPlan<T>::solve(bool withPhi, const int* skip) {
    if (fft_) {
        solveFft(withPhi, skip);
        return;
    }
    const int M = mx_, N = my_;
    SpectralEpi<T> none{};
    // Intermediates alternate between transposed and plain layouts so that
    // every product reads both operands along k: T0^T, P, T1^T / T2^T, then
    // the fields in the plain [x][y] layout.
    {   // T0^T = (rho * Cy^T)^T
        GemmBatch<T> gb{};
        gb.A[0] = rho_; gb.B[0] = cyT_; gb.C[0] = t0_;
        gemmDispatch<T, kEpiNone, true>(1, gb, M, N, N, N, N, M, none, stream_, skip, 1, x3_);
    }
    {   // P = spectral(Cx * T0)
        GemmBatch<T> gb{};
        gb.A[0] = cx_; gb.B[0] = t0_; gb.C[0] = p_;
        SpectralEpi<T> epi{u_, v_, nx_, ny_, invNorm_};
        gemmDispatch<T, kEpiSpectral, false>(1, gb, M, N, M, M, M, N, epi, stream_, skip, 1,
                                             x3_);
    }
    {   // T1^T = (P * Ay^T)^T, T2^T = (P * Bsv)^T
        GemmBatch<T> gb{};
        gb.A[0] = p_; gb.B[0] = ayT_; gb.C[0] = t1_;
        gb.A[1] = p_; gb.B[1] = bsv_; gb.C[1] = t2_;
        gemmDispatch<T, kEpiNone, true>(2, gb, M, N, N, N, N, M, none, stream_, skip, 1, x3_);
    }
    {   // xi_x = Sxu * T1, xi_y = Ax * T2 (interleaved or not), phi = Ax * T1
        GemmBatch<T> gb{};
        gb.A[0] = sxu_; gb.B[0] = t1_; gb.C[0] = interleaved_ ? exy_ : ex_;
        gb.A[1] = ax_;  gb.B[1] = t2_; gb.C[1] = interleaved_ ? exy_ + 1 : ey_;
        gemmDispatch<T, kEpiNone, false>(2, gb, M, N, M, M, M, N, none, stream_, skip,
                                         interleaved_ ? 2 : 1, x3_);
        if (withPhi) {
            GemmBatch<T> gp{};
            gp.A[0] = ax_; gp.B[0] = t1_; gp.C[0] = phi_;
            gemmDispatch<T, kEpiNone, false>(1, gp, M, N, M, M, M, N, none, stream_, skip, 1,
                                             x3_);
        }
    }
}

// ------------------------------------------------------------ no allocation
// The factor tables of one axis: the twCount(L) entries of exp(-2 pi i t/4L)
// (Twid; complex, interleaved, FP64), the values PoissonPlan uploads.
__global__ __launch_bounds__(kBlock) void k_fft_tables(double* tab, int L) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const double pi = 3.14159265358979323846;
    if (k < twCount(L)) {
        const int nLo = twLo(L);
        const double e = k < nLo ? k : static_cast<double>((k - nLo) << twShift(L));
        tab[2 * k] = cos(2 * pi * e / (4.0 * L));
        tab[2 * k + 1] = -sin(2 * pi * e / (4.0 * L));
    }
}

size_t fftSolveWorkFloats(int nx, int ny) {
    return 3 * static_cast<size_t>(nx) * ny + 5 * static_cast<size_t>(nx + ny);
}

void fftSolveInto(int nx, int ny, float stepX, float stepY, const float* rho, float* out3,
                  float* work, hipStream_t s) {
    if (!PoissonPlan<float>::fftShape(nx) || !PoissonPlan<float>::fftShape(ny) || nx < 64 ||
        ny < 64)
        throw std::invalid_argument("fftSolveInto: grid sides must be powers of two, 64 to 1024");
    const size_t n = static_cast<size_t>(nx) * ny;
    float* t0 = work;
    float* t1 = t0 + n;
    float* t2 = t1 + n;
    // each axis's FP64 factor tables in the 4 n floats the work array gives it
    double* tabX = reinterpret_cast<double*>(t2 + n);
    double* tabY = reinterpret_cast<double*>(t2 + n + 4 * static_cast<size_t>(nx));
    static_assert(4 * twCount(64) <= 4 * 64 && 4 * twCount(1024) <= 4 * 1024,
                  "an axis's factor tables fit the 4 n floats the work array gives them");
    launchKernel(k_fft_tables, dim3((twCount(nx) + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                 tabX, nx);
    launchKernel(k_fft_tables, dim3((twCount(ny) + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                 tabY, ny);
    SPHIP_CHECK(hipGetLastError());
    const FftTab<float> tx{reinterpret_cast<const cplx<double>*>(tabX), nx, stepX};
    const FftTab<float> ty{reinterpret_cast<const cplx<double>*>(tabY), ny, stepY};
    // the LDS of the plan's passes with one column pair per block (setFftLds)
    const size_t c = 2 * sizeof(float);
    const size_t ldsA = (kFftLdsHead<float>(ny) + 2 * static_cast<size_t>(ny)) * c;
    const size_t ldsB = (kFftLdsHead<float>(nx) + 3 * static_cast<size_t>(nx)) * c;
    const size_t ldsC = (kFftLdsHead<float>(ny) + 4 * static_cast<size_t>(ny)) * c;
    raiseLdsLimit(reinterpret_cast<const void*>(k_dct_rows<float>), ldsA);
    raiseLdsLimit(reinterpret_cast<const void*>(k_dct_cols<float, 256>), ldsB);
    raiseLdsLimit(reinterpret_cast<const void*>(k_idct_rows<float>), ldsC);
    const float invNorm = static_cast<float>(1.0 / (4.0 * nx * ny));
    launchKernel(k_dct_rows<float>, dim3(nx / 2), dim3(fftThreads(ny)), static_cast<uint32_t>(ldsA),
                 s, rho, t0, nx, ny, ny, ty, static_cast<const int*>(nullptr));
    launchKernel(k_dct_cols<float, 256>, dim3(ny / 2), dim3(fftThreads(nx)), static_cast<uint32_t>(ldsB),
                 s, t0, t1, t2, 1, nx, ny, ny, tx, ty, invNorm, static_cast<const int*>(nullptr));
    launchKernel(k_idct_rows<float>, dim3(nx), dim3(fftThreads(ny)), static_cast<uint32_t>(ldsC),
                 s, t1, t2, out3 + n, out3 + 2 * n, out3, 1, nx, ny, ny, ty,
                 static_cast<const int*>(nullptr));
    SPHIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ batches
namespace {

template <typename T, int PASS>
__global__ __launch_bounds__(kBlock) void k_fft_batch(const PoissonBatchArgs<T>* __restrict__ args,
                                                      const int4* __restrict__ map, int nx, int ny,
                                                      int ld) {
    const int4 m = map[blockIdx.x];
    const PoissonBatchArgs<T>& a = args[m.x];
    if (a.skip && *a.skip) return;
    const FftTab<T> tx{reinterpret_cast<const cplx<double>*>(a.tabX), nx, a.stepX};
    const FftTab<T> ty{reinterpret_cast<const cplx<double>*>(a.tabY), ny, a.stepY};
    if constexpr (PASS == 0)
        dctRowsBlock<T>(a.rho, a.t0, nx, ny, ld, ty, m.y);
    else if constexpr (PASS == 1)
        dctColsBlock<T>(a.t0, a.t1, a.t2, 1, nx, ny, ld, tx, ty, a.invNorm,
                              xcdBlock(m.y, m.z));
    else
        idctRowsBlock<T>(a.t1, a.t2, a.exy, a.exy + 1, static_cast<T*>(nullptr), 2, nx, ny, ld, ty,
                         m.y);
}

}  // namespace

template <typename T>
PoissonBatchArgs<T> PoissonPlan<T>::batchArgs(const int* skip) const {
    if (!fft_ || !interleaved_) throw std::logic_error("batched solve: FFT plan with interleaved field only");
    return PoissonBatchArgs<T>{rho_, t0_, t1_, t2_, exy_, tabX_, tabY_, stepX_, stepY_, invNorm_,
                               skip};
}

template <typename T>
void launchPoissonBatch(int pass, const PoissonBatchArgs<T>* args, const int4* map, int blocks,
                        int nx, int ny, int ld, hipStream_t s) {
    if (blocks <= 0) return;
    const size_t c = 2 * sizeof(T);
    // the plan's LDS with one column pair per block (setFftLds), the inverse
    // rows without phi
    const size_t ldsA = (kFftLdsHead<T>(ny) + 2 * static_cast<size_t>(ny)) * c;
    const size_t ldsB = (kFftLdsHead<T>(nx) + 3 * static_cast<size_t>(nx)) * c;
    const size_t ldsC = (kFftLdsHead<T>(ny) + 2 * static_cast<size_t>(ny)) * c;
    static std::once_flag once;
    std::call_once(once, [] {
        // the largest grid's LDS (1024 per side) covers every smaller one
        const size_t big = (kFftLdsHead<T>(1024) + 4 * 1024) * 2 * sizeof(T);
        raiseLdsLimit(reinterpret_cast<const void*>(k_fft_batch<T, 0>), big);
        raiseLdsLimit(reinterpret_cast<const void*>(k_fft_batch<T, 1>), big);
        raiseLdsLimit(reinterpret_cast<const void*>(k_fft_batch<T, 2>), big);
    });
    if (pass == 0)
        launchKernel(k_fft_batch<T, 0>, dim3(blocks), dim3(fftThreads(ny)),
                     static_cast<uint32_t>(ldsA), s, args, map, nx, ny, ld);
    else if (pass == 1)
        launchKernel(k_fft_batch<T, 1>, dim3(blocks), dim3(fftThreads(nx)),
                     static_cast<uint32_t>(ldsB), s, args, map, nx, ny, ld);
    else
        launchKernel(k_fft_batch<T, 2>, dim3(blocks), dim3(fftThreads(ny)),
                     static_cast<uint32_t>(ldsC), s, args, map, nx, ny, ld);
    SPHIP_CHECK(hipGetLastError());
}

template class PoissonPlan<float>;
template class PoissonPlan<double>;
template void launchPoissonBatch<float>(int, const PoissonBatchArgs<float>*, const int4*, int, int,
                                        int, int, hipStream_t);
template void launchGemm<float>(int, const float* const*, const float* const*, float* const*,
                                int, int, int, int, int, int, hipStream_t, bool);
template void launchGemm<double>(int, const double* const*, const double* const*,
                                 double* const*, int, int, int, int, int, int, hipStream_t, bool);

}  // namespace sphip
