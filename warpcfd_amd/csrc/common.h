// Shared helpers for the wcfd HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/wcfd.h"

namespace wcfd {

constexpr int kBlock = 256;           // threads per workgroup for streaming kernels
constexpr int kWave = 64;             // CDNA wavefront
constexpr int kMaxWaves = 16;         // LDS sizing: workgroups of up to 1024 threads
constexpr int kMaxReduceBlocks = 2048;  // upper bound on partials of one reduction
constexpr int kKrylovGrid = 1024;       // workgroups of the Krylov kernels (grid-stride)

// per-thread HIP error of the last failing call
extern thread_local int g_last_hip_error;

// --------------------------------------------------------- opt-in profiler
// Every launch site is bracketed by a ProfScope.  While wcfd_profile_begin()
// is active each launch records a start and a stop hipEvent on its own
// stream; wcfd_profile_end() folds them into per-kernel totals.
enum KernelId {
  K_SET_BC, K_FACE_INTERP, K_GREEN_GAUSS, K_MASS_FLUX, K_DIVERGENCE, K_ASM_MOM, K_ASM_P,
  K_RUA_HBYA, K_RELAX_P, K_FACE_P_FC, K_GRAD_P_CORR, K_INV_DIAG, K_SPMV, K_DOT, K_CG_INIT,
  K_CG_1, K_CG_2, K_BI_INIT, K_BI_1, K_BI_2, K_BI_3, K_FOLD, K_ELL_TO_CSR, K_COUNT
};
bool prof_active();
void prof_record(int id, hipStream_t s, bool start);
struct ProfScope {
  int id;
  hipStream_t s;
  bool on;
  ProfScope(int id_, hipStream_t s_) : id(id_), s(s_), on(prof_active()) {
    if (on) prof_record(id, s, true);
  }
  ~ProfScope() {
    if (on) prof_record(id, s, false);
  }
};

// launch KERNEL<<<GRID, kBlock, 0, S>>>(args...) inside a profiler scope
#define WCFD_LAUNCH(ID, S, KERNEL, GRID, ...)          \
  do {                                                \
    ::wcfd::ProfScope wcfd_ps_((ID), (S));            \
    KERNEL<<<(GRID), kBlock, 0, (S)>>>(__VA_ARGS__);  \
  } while (0)

// the same with a run-time workgroup size
#define WCFD_LAUNCH_B(ID, S, KERNEL, GRID, BLOCK, ...)   \
  do {                                                  \
    ::wcfd::ProfScope wcfd_ps_((ID), (S));              \
    KERNEL<<<(GRID), (BLOCK), 0, (S)>>>(__VA_ARGS__);   \
  } while (0)

// Launch shape of the Krylov kernels (workgroups, threads per workgroup),
// set by the autotuner through wcfd_set_krylov_launch(); process-wide.
struct KrylovLaunch {
  int grid;
  int block;
};
KrylovLaunch krylov_launch();

inline int check_hip(hipError_t e) {
  if (e != hipSuccess) {
    g_last_hip_error = static_cast<int>(e);
    return WCFD_EHIP;
  }
  return WCFD_OK;
}

inline int check_launch() { return check_hip(hipGetLastError()); }

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return static_cast<unsigned>(g < 1 ? 1 : g);
}

inline unsigned reduce_grid(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  if (g > kMaxReduceBlocks) g = kMaxReduceBlocks;
  return static_cast<unsigned>(g < 1 ? 1 : g);
}

inline size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Carves a caller-provided workspace into aligned pieces.
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* p) : base(static_cast<char*>(p)) {}
  template <class T>
  T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
    off += align256(count * sizeof(T));
    return p;
  }
};

// ------------------------------------------------------ XCD-aware mapping
// MI355X deals workgroups to its 8 XCDs round-robin (workgroup b runs on XCD
// b % 8) and each XCD has its own L2.  xcd_remap() renumbers workgroups so
// that the ones on one XCD get consecutive virtual ids, i.e. one contiguous
// band of rows / faces per XCD: the neighbour gathers of a band (RCM keeps
// them within the band width) then hit that XCD's L2 instead of another's.
// It is a bijection on [0, G) for any G.
constexpr unsigned kNumXcd = 8;
__device__ inline unsigned xcd_remap(unsigned b, unsigned G) {
  const unsigned x = b % kNumXcd, q = G / kNumXcd, r = G % kNumXcd;
  return x * q + (x < r ? x : r) + b / kNumXcd;
}

// Item index of a one-item-per-thread launch with the XCD-aware mapping.
__device__ inline int xcd_item() {
  return (int)(xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
}

// Rows of a reduction kernel: workgroup (virtual id vb) owns the contiguous
// chunk [vb*chunk, (vb+1)*chunk) and its threads stride through it.
struct RowRange {
  int begin, end;
};
__device__ inline RowRange my_rows(int n) {
  const int vb = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int chunk = (n + (int)gridDim.x - 1) / (int)gridDim.x;
  const int b = vb * chunk;
  return {b, b + chunk < n ? b + chunk : n};
}

// ------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block sum of ND values per thread; result valid in thread 0.  Order is fixed
// by (blockDim, thread index), so the result is deterministic.
template <int ND>
__device__ inline void block_sum(double (&v)[ND], double* lds /* [ND][kMaxWaves] */) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    double s = wave_sum(v[i]);
    if (lane == 0) lds[i * kMaxWaves + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += lds[i * kMaxWaves + w];
      v[i] = s;
    }
  }
  __syncthreads();
}

// ------------------------------------------------------- matrix access
// Krylov kernels are written once against these two row views.
//
// CsrMat: a general CSR matrix (row_ptr, col, values, diagonal position).
struct CsrMat {
  int n;
  const int32_t* rp;
  const int32_t* col;
  const float* val;
  const int32_t* dpos;
  __device__ float diag(int i) const { return val[dpos[i]]; }
  template <class F>
  __device__ void row(int i, F&& f) const {
    for (int j = rp[i]; j < rp[i + 1]; ++j) f(val[j], col[j]);
  }
};

// EllMat: the device path's matrices, slot-major.  val[i] is the diagonal of
// row i and val[(k+1)*n + i] the coefficient of the neighbour in local face
// slot k (0 at boundary slots, whose neighbour index is the row itself).
// Every load is coalesced across a wavefront and no row pointer is read.
// With D16 the neighbour is stored as an int16 offset from the row (the RCM
// band fits), halving the index bytes.
template <bool D16>
struct EllMat {
  int n;
  const float* val;
  const int32_t* nbr;
  const int16_t* dnb;
  __device__ float diag(int i) const { return val[i]; }
  template <class F>
  __device__ void row(int i, F&& f) const {
    f(val[i], i);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int c = D16 ? i + (int)dnb[k * n + i] : nbr[k * n + i];
      f(val[(k + 1) * n + i], c);
    }
  }
};

}  // namespace wcfd
