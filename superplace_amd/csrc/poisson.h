// DCT Poisson solve (host-side plan): LDS-resident FFTs for power-of-two
// grids, batched dense transforms on MFMA otherwise.
//
// Solves lap(phi) = -rho with Neumann boundaries and returns the field
// xi = -grad(phi) with the conventions of the reference's CPU backend
// (src/numeric/cpu/cpu_poisson.cpp:3-40, :149-207): forward REDFT10 on both
// axes, P = F / (4 nx ny) / (u_j^2 + v_k^2) with P[0][0] = 0, then
//   phi  = REDFT01_x REDFT01_y P
//   xi_x = RODFT01_x REDFT01_y (P shifted down one row, rows scaled by u)
//   xi_y = REDFT01_x RODFT01_y (P shifted left one column, columns scaled by v)
// Power-of-two grids run three FFT passes (rows, columns, rows).  Other
// sizes run each 1-D transform along an axis as a product with a dense
// nx x nx (or ny x ny) matrix; the index shift and the u/v scaling are folded
// into the sine matrices, and the spectral divide is the epilogue of the
// second forward product.  Layout: row-major, x slow, y fast (data[x * ld + y]),
// padded to a multiple of 64 on both axes with zero rows/columns.
#pragma once

#include "common.h"
#include "devmem.h"

namespace sphip {

template <typename T>
struct WaTask;
template <typename T>
struct DensityIn;

// LDS bound of the column pass that carries the WA blocks (all blocks of a
// launch reserve it, so a larger one starves the WA blocks of occupancy).
constexpr size_t kFusedLdsMax = 32 * 1024;

// One instance's buffers and tables for a batched launch of the FFT passes
// (instances of the same grid size; see launchPoissonBatch).
template <typename T>
struct PoissonBatchArgs {
    const T* rho;
    T* t0;
    T* t1;
    T* t2;
    T* exy;             // interleaved field out
    const double* tabX;   // per axis: the FP64 twiddle factor tables (complex)
    const double* tabY;
    double stepX, stepY;   // bin steps (the spectral scales are computed from them)
    T invNorm;
    const int* skip;    // the instance's done flag
};

// PoissonPlan options: kPlanDense runs the dense transforms on every grid
// (power-of-two grids too, which otherwise take the FFT passes); kPlanBf16x3
// runs the dense products as split-BF16 products (FP32 plans only).
enum : unsigned { kPlanDense = 1u, kPlanBf16x3 = 2u };

template <typename T>
class PoissonPlan {
public:
    // interleaved: the field is written as (xi_x, xi_y) pairs into fieldXY()
    // (element 2 (x * ld + y) + axis) instead of fieldX() / fieldY().
    // arena: the plan's arrays are reserved there (and the arena owns them);
    // they exist once the caller has placed the arena.  Without, the plan
    // allocates them and is ready on return.  opts: kPlan* bits.
    PoissonPlan(int nx, int ny, double stepX, double stepY, hipStream_t stream,
                bool interleaved = false, DevArena* arena = nullptr, unsigned opts = 0);
    ~PoissonPlan();
    PoissonPlan(const PoissonPlan&) = delete;
    PoissonPlan& operator=(const PoissonPlan&) = delete;

    // rho() must hold the padded density (zeros in the padding).  Writes
    // fieldX()/fieldY() and, if withPhi, phi().  Enqueued on the stream; the
    // kernels return at once while *skip (a device flag, may be null) is set.
    void solve(bool withPhi, const int* skip = nullptr);
    // solve(false, skip) with the WA pass's waBlocks blocks carried by the
    // column pass's launch (the same partials as launching the WA pass
    // alone with waBlocks blocks).  Returns false, enqueuing nothing, where
    // the grid does not allow it (dense path, or LDS above kFusedLdsMax).
    // With din, the first pass finalizes the density grids itself (rho is
    // then not written): one tau partial per block of that pass into
    // din->partials, their count into *densityBlocks.
    // WA blocks that fit beside the column blocks in the fused launch (the
    // blocks the GPU holds at once, less the column blocks); 0 when the plan
    // has no fused launch
    int fusedWaResidentBlocks() const;
    bool solveWithWa(const WaTask<T>& wa, int waBlocks, const int* skip,
                     const DensityIn<T>* din = nullptr, int* densityBlocks = nullptr);
    // solve(false, skip) from the integer density grids (takesDensity()
    // plans only); returns the tau partial count.
    int solveFromDensity(const DensityIn<T>& din, const int* skip);
    bool takesDensity() const;

    int nx() const { return nx_; }
    int ny() const { return ny_; }
    int ld() const { return my_; }          // row stride of every grid buffer
    int rowsPadded() const { return mx_; }
    T* rho() { return rho_; }
    T* phi() { return phi_; }
    T* fieldX() { return ex_; }
    T* fieldY() { return ey_; }
    T* fieldXY() { return exy_; }
    hipStream_t stream() const { return stream_; }
    // Floating-point operations of one solve (dense products actually
    // executed; 0 on the FFT path).
    double flops(bool withPhi) const;
    // Grid bytes one solve moves through memory on the FFT path (0 on the
    // dense path).
    double bytes(bool withPhi) const;
    // Power-of-two grids (4 .. 1024 per side) take the FFT path.
    bool usesFft() const { return fft_; }
    bool usesBf16x3() const { return x3_; }
    // The plan's buffers for a batched solve (interleaved FFT plans only).
    PoissonBatchArgs<T> batchArgs(const int* skip) const;
    static bool fftShape(int n);

private:
    int nx_, ny_, mx_, my_;
    hipStream_t stream_;
    bool interleaved_;
    bool ownsMem_ = true;
    T *rho_ = nullptr, *phi_ = nullptr, *ex_ = nullptr, *ey_ = nullptr, *exy_ = nullptr;
    T *t0_ = nullptr, *p_ = nullptr, *t1_ = nullptr, *t2_ = nullptr;
    T *cyT_ = nullptr, *cx_ = nullptr, *ayT_ = nullptr, *bsv_ = nullptr, *sxu_ = nullptr,
      *ax_ = nullptr;
    T *u_ = nullptr, *v_ = nullptr;
    T invNorm_;
    double stepX_, stepY_;
    // FFT path: per axis the twiddle factor tables (complex, interleaved);
    // LDS bytes of the three passes
    bool fft_ = false;
    bool x3_ = false;   // dense products as split-BF16 products
    double *tabX_ = nullptr, *tabY_ = nullptr;
    size_t ldsA_ = 0, ldsB_ = 0, ldsC_ = 0, ldsBW_ = 0;
    int colPairs_ = 1;   // column pairs per block of the column pass
    void setFftLds();

public:
    // Column pairs per block of the FFT column pass when it runs on its own
    // (a launch shape: every pair is transformed alike, so the results do
    // not depend on it).  False, with the plan unchanged, where the pairs
    // do not divide the columns or their LDS would exceed 150 KB.
    bool setColumnPairs(int pairs);
    int columnPairs() const { return colPairs_; }
    void solveFft(bool withPhi, const int* skip);
    int passA(const DensityIn<T>* din, const int* skip);   // returns the tau partial count
    void passBC(const int* skip, bool withPhi = false);
};

// The FP32 FFT solve of an nx x ny grid (powers of two, 64 to 1024 per side,
// no padding) over caller-provided device memory, allocating nothing: rho in
// (nx * ny), out3 = phi | xi_x | xi_y (3 nx ny), and `work` of
// fftSolveWorkFloats(nx, ny) floats for the intermediates, the tables and
// the spectral scales (rewritten by every call).  The kernels and their
// arithmetic are PoissonPlan<float>'s with one column pair per block, so the
// results equal a plan's solve(true).  Enqueued on s.
size_t fftSolveWorkFloats(int nx, int ny);
void fftSolveInto(int nx, int ny, float stepX, float stepY, const float* rho, float* out3,
                  float* work, hipStream_t s);

// Pass `pass` (0: rows, 1: columns with the spectral step, 2: inverse rows
// into the interleaved field) of the FFT solve for many instances of one
// nx x ny grid (row stride ld) in one launch: block b runs block map[b].y of
// instance map[b].x's pass (of map[b].z blocks), exactly as the plan's own
// launch of that pass with one column pair per block does.  `blocks` is
// the total; blocks() gives an instance's count per pass.
template <typename T>
void launchPoissonBatch(int pass, const PoissonBatchArgs<T>* args, const int4* map, int blocks,
                        int nx, int ny, int ld, hipStream_t s);
inline int poissonBatchBlocks(int pass, int nx, int ny) {
    return pass == 0 ? nx / 2 : (pass == 1 ? ny / 2 : nx);
}

// Batched C = A * B^T (A M x K, B N x K, row-major, dims multiples of 64) on
// MFMA; the plan's primitive, exposed for the GEMM unit test.  x3: as
// split-BF16 products (T = float only).
template <typename T>
void launchGemm(int batch, const T* const* A, const T* const* B, T* const* C, int M, int N,
                int K, int lda, int ldb, int ldc, hipStream_t stream, bool x3 = false);

}  // namespace sphip
