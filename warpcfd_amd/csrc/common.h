// Shared helpers for the wcfd HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/wcfd.h"

namespace wcfd {

// Field type of this build (float, or double for libwcfd64.so).
using real = wcfd_real;
static_assert(sizeof(real) == 4 || sizeof(real) == 8, "real is float or double");

constexpr int kBlock = 256;           // threads per workgroup for streaming kernels
constexpr int kWave = 64;             // CDNA wavefront
constexpr int kMaxWaves = 16;         // LDS sizing: workgroups of up to 1024 threads
constexpr int kMaxReduceBlocks = 2048;  // upper bound on partials of one reduction
constexpr int kKrylovGrid = 1024;       // workgroups of the Krylov kernels (grid-stride)
constexpr int kMaxKrylovBlocks = 65536; // upper bound on the Krylov kernels' workgroups
constexpr int kMaxInlineFold = 2048;    // default: largest per-case fold done inline
// build-time cache hints of the Krylov row kernels: nontemporal stores of
// the per-row outputs (on: 4-5% faster CG iterations at 0.3M and 4.4M
// cells on MI355X), nontemporal loads of the matrix stream (off: faster
// only once the matrix no longer fits the 256 MB Infinity Cache, and slower
// together with the nontemporal stores)
#ifndef WCFD_NT_STORE
#define WCFD_NT_STORE 1
#endif
// face interpolation with the owners' data staged in LDS (k_face_interp_lds)
// or gathered (k_face_interp, the default); the two agree bit for bit.  The
// LDS form measured 18-22 % slower at 0.32-8.6 M cells: its barrier makes
// the workgroup wait for its slowest load, and the L1 already serves the
// owners' reuse.
#ifndef WCFD_FACE_LDS
#define WCFD_FACE_LDS 0
#endif
#ifndef WCFD_NT_MATRIX
#define WCFD_NT_MATRIX 0
#endif
// The cell kernels that gather face data through their slots (assembly,
// the pressure correction) read and write their cell-ordered streams (slot
// tables, the cell's own values, the outputs) with the nontemporal hint, so
// that these do not push out of the L2 the face data that a face's
// neighbour side reads again one band of cells later.  Measured at 8.64 M
// cells: FETCH_SIZE of the momentum assembly 2281 -> 1782 MB, 2 % faster.
#ifndef WCFD_NT_CELL
#define WCFD_NT_CELL 1
#endif
// the same for the face-ordered streams of the internal faces in the face
// kernels (indices, face geometry, the face outputs), which keeps the
// gathered cell data in the L2: face interpolation at 8.64 M cells 3.6 %
// faster, FETCH_SIZE 1640 -> 1488 MB
#ifndef WCFD_NT_FACE
#define WCFD_NT_FACE 1
#endif
// Largest workgroup of the row-parallel Krylov / reduction kernels: their
// __launch_bounds__, which lets a wave hold up to 256 VGPRs (a whole row of
// gathers in flight) without spilling.
constexpr int kMaxKrylovBlock = 512;

// the mesh descriptor's slot count: 0 (six), 4, 5 or 6
inline bool slots_ok(const wcfd_mesh* m) {
  return m->n_slots == 0 || (m->n_slots >= 4 && m->n_slots <= 6);
}

// per-thread HIP error of the last failing call
extern thread_local int g_last_hip_error;

// --------------------------------------------------------- opt-in profiler
// While wcfd_profile_begin() is active every launch takes a start and a stop
// hipEvent that the runtime stamps with the kernel's own begin and end
// (hipExtLaunchKernelGGL), so a kernel's time excludes the launch gap before
// it; wcfd_profile_end() folds them into per-kernel totals.  Otherwise
// kernels launch with <<<>>> and no event.
enum KernelId {
  K_SET_BC, K_FACE_INTERP, K_GREEN_GAUSS, K_MASS_FLUX, K_DIVERGENCE, K_ASM_MOM, K_ASM_P,
  K_RUA_HBYA, K_RELAX_P, K_FACE_P_FC, K_GRAD_P_CORR, K_INV_DIAG, K_SPMV, K_DOT, K_CG_INIT,
  K_CG_1, K_CG_2, K_BI_INIT, K_BI_1, K_BI_2, K_BI_3, K_FOLD, K_ELL_TO_CSR, K_TERM_WEIGHTS,
  K_EQUATION, K_CONVERGENCE, K_CG_XCD, K_COUNT
};
bool prof_active();
// a start/stop event pair for one launch of kernel `id` (false: not
// profiling, or the event pool is used up)
bool prof_events(int id, hipEvent_t* start, hipEvent_t* stop);
// Krylov iteration index attached to the launches that follow on this host
// thread (-1: not an iteration launch)
void prof_set_iter(int it);
// position in the launch records, and retirement of the records from there
// on whose iteration index is >= iters_run (launches that exited at entry)
size_t prof_mark();
void prof_retire(size_t from, int iters_run);

// launch KERNEL<<<GRID, BLOCK, 0, S>>>(args...), timed when profiling
#define WCFD_LAUNCH_B(ID, S, KERNEL, GRID, BLOCK, ...)                                   \
  do {                                                                                  \
    hipEvent_t wcfd_e0_ = nullptr, wcfd_e1_ = nullptr;                                  \
    if (::wcfd::prof_active() && ::wcfd::prof_events((ID), &wcfd_e0_, &wcfd_e1_))       \
      hipExtLaunchKernelGGL((KERNEL), dim3(GRID), dim3(BLOCK), 0, (S), wcfd_e0_, wcfd_e1_, \
                            0u, __VA_ARGS__);                                           \
    else                                                                                \
      KERNEL<<<(GRID), (BLOCK), 0, (S)>>>(__VA_ARGS__);                                 \
  } while (0)

// the same with kBlock threads per workgroup
#define WCFD_LAUNCH(ID, S, KERNEL, GRID, ...) WCFD_LAUNCH_B(ID, S, KERNEL, GRID, kBlock, __VA_ARGS__)

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

// Workgroups of a one-row-per-thread reduction over n cells of kBlock rows,
// at least kMaxReduceBlocks (the partial rows a workspace holds).
inline int64_t cell_reduce_blocks(int64_t n) {
  const int64_t g = (n + kBlock - 1) / kBlock;
  return g > kMaxReduceBlocks ? g : kMaxReduceBlocks;
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

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// wave_sum of ND values at once: each butterfly step issues the exchanges of
// all ND values before adding, so their LDS-crossbar latencies overlap (one
// wait per step instead of one per value and step).  Every value is summed in
// wave_sum's order, so the results are bitwise the same.
template <int ND>
__device__ inline void wave_sum_n(double (&v)[ND]) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    double o[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) o[i] = __shfl_xor(v[i], off, kWave);
#pragma unroll
    for (int i = 0; i < ND; ++i) v[i] += o[i];
  }
}

// ------------------------------------------------ uniform base + offset
// base (uniform across the wave: a kernel argument plus uniform terms) plus
// a 32-bit byte offset per lane.  The compiler forms such an address as a
// scalar base plus one offset VGPR (global_load ... vOff, s[base]), so the
// loads of several arrays at the same row share one offset register instead
// of a 64-bit address pair each.  Offsets must stay below 2^32 bytes:
// mesh_ok() refuses meshes whose field arrays reach that.
//
// The kernels form every offset of a field array as index * 4 -- the byte
// offset of a 4-byte element, shared with the int32 tables at the same
// index.  An 8-byte element type (real = double in libwcfd64.so) scales it
// by 2 here; every other element type takes the offset as a byte offset.
template <class T>
__device__ __forceinline__ uint32_t scaled(uint32_t byte_off) {
  if constexpr (sizeof(T) == 8)
    return byte_off * 2u;
  else
    return byte_off;
}
template <class T>
__device__ __forceinline__ T ld_off(const T* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + scaled<T>(byte_off));
}
template <class T>
__device__ __forceinline__ void st_off(T* base, uint32_t byte_off, T v) {
#if WCFD_NT_STORE
  __builtin_nontemporal_store(
      v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + scaled<T>(byte_off)));
#else
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + scaled<T>(byte_off)) = v;
#endif
}
// an ordinary store at a uniform base plus a 32-bit byte offset
template <class T>
__device__ __forceinline__ void st_plain(T* base, uint32_t byte_off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + scaled<T>(byte_off)) = v;
}
// a cell-ordered array read or written once in a launch, with the
// nontemporal hint when NT
template <bool NT, class T>
__device__ __forceinline__ T ld_cell(const T* base, uint32_t byte_off) {
  if constexpr (NT)
    return __builtin_nontemporal_load(
        reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + scaled<T>(byte_off)));
  else
    return ld_off(base, byte_off);
}
template <bool NT, class T>
__device__ __forceinline__ void st_cell(T* base, uint32_t byte_off, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(
        v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + scaled<T>(byte_off)));
  else
    st_plain(base, byte_off, v);
}
// the same at a plain pointer
template <bool NT, class T>
__device__ __forceinline__ T ld_once(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, class T>
__device__ __forceinline__ void st_once(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
// matrix stream (read once per iteration)
template <class T>
__device__ __forceinline__ T ld_mat(const T* base, uint32_t byte_off) {
#if WCFD_NT_MATRIX
  return __builtin_nontemporal_load(
      reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + scaled<T>(byte_off)));
#else
  return ld_off(base, byte_off);
#endif
}

// Raw buffer view of a real array for gathers: the descriptor sits in
// scalar registers and each load takes one 32-bit byte offset per lane
// (buffer_load_dword vOff, s[desc]), so the gathers of several vectors at
// the same neighbour share one offset register.  Offsets at or beyond
// `bytes` read 0.  Word 3 is the gfx9 raw dword-buffer format.  Sizes and
// offsets are given for 4-byte elements (index * 4) and scaled as above.
constexpr int kBufWord3 = 0x00020000;
struct FBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ FBuf(const real* p, uint32_t bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<real*>(p), (short)0,
                                            (int)scaled<real>(bytes), kBufWord3)) {}
  __device__ __forceinline__ real operator()(uint32_t off) const { return load<real>(off); }
  template <class T>
  __device__ __forceinline__ T load(uint32_t off) const {
    if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, scaled<T>(off), 0, 0));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  }
};

// a * b + c rounded once, in the field type: kernels that must round alike
// spell their multiply-adds out instead of leaving them to contraction
__device__ __forceinline__ float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double mad(double a, double b, double c) { return __builtin_fma(a, b, c); }

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

// ------------------------------------------------------- segmentation
// Rows of a reduction kernel.  The rows [0, n) hold one or more independent
// cases; case k owns rows [cell_off[k], cell_end[k]) and virtual
// workgroups [blk_off[k], blk_off[k+1]).  Virtual workgroup vb (the XCD
// remap of blockIdx.x) owns the contiguous chunk of `chunk` rows starting at
// cell_off[k] + (vb - blk_off[k]) * chunk of its case, and its threads
// stride through it.  With cell_off == NULL there is one case and
// workgroup vb owns [vb*chunk, (vb+1)*chunk).  Because a case's chunks,
// partials and fold order depend only on the case itself (and the chunk and
// workgroup size), a case gives bitwise the same sums alone or in a batch.
struct Seg {
  int n;                    // rows in total
  int G;                    // virtual workgroups in total
  int chunk;                // rows per workgroup
  int n_cases;              // K
  const int32_t* cell_off;  // [K+1]; NULL: one case owning [0, n)
  const int32_t* cell_end;  // [K]; NULL: cell_off[k+1]
  const int32_t* blk_off;   // [K+1]
  const int32_t* blk_case;  // [G]
  const uint8_t* active;    // [K]; NULL: every case active
  // 1: the partials the Krylov iteration kernels write are reduced to one
  // value per case (at the case's first slot) by a fold kernel before the
  // next kernel reads them; 0: every reader folds all of its case's partials
  int prefold;
};

struct BlockRows {
  int vb, k, begin, end, b0, nb;
  int nf;   // partials a reader of iteration partials folds: nb, or 1 when prefolded
};

__device__ inline BlockRows block_rows(const Seg& s) {
  BlockRows r;
  r.vb = (int)xcd_remap(blockIdx.x, gridDim.x);
  if (!s.cell_off) {
    r.k = 0;
    r.b0 = 0;
    r.nb = s.G;
    r.begin = r.vb * s.chunk;
    r.end = r.begin + s.chunk < s.n ? r.begin + s.chunk : s.n;
  } else {
    r.k = s.blk_case[r.vb];
    r.b0 = s.blk_off[r.k];
    r.nb = s.blk_off[r.k + 1] - r.b0;
    r.begin = s.cell_off[r.k] + (r.vb - r.b0) * s.chunk;
    const int e = s.cell_end ? s.cell_end[r.k] : s.cell_off[r.k + 1];
    r.end = r.begin + s.chunk < e ? r.begin + s.chunk : e;
  }
  r.nf = s.prefold ? 1 : r.nb;
  return r;
}

// first virtual workgroup and workgroup count of case k
__device__ inline void case_blocks(const Seg& s, int k, int& b0, int& nb) {
  b0 = s.blk_off ? s.blk_off[k] : 0;
  nb = s.blk_off ? s.blk_off[k + 1] - b0 : s.G;
}

// Single-case segmentation of n rows over at most max_g workgroups of
// `block` threads.
inline Seg single_seg(int n, int block, int max_g) {
  int64_t g = (n + (int64_t)block - 1) / block;
  if (g > max_g) g = max_g;
  if (g < 1) g = 1;
  Seg sg{};
  sg.n = n;
  sg.G = (int)g;
  sg.chunk = (int)((n + g - 1) / g);
  sg.n_cases = 1;
  return sg;
}

// Segmentation of a batch descriptor (rows = cells of the batch mesh).
inline Seg batch_seg(const wcfd_batch& bt, int n) {
  Seg sg{};
  sg.n = n;
  sg.G = bt.n_blocks;
  sg.chunk = bt.chunk;
  sg.n_cases = bt.n_cases;
  sg.cell_off = bt.cell_off;
  sg.cell_end = bt.cell_end;
  sg.blk_off = bt.blk_off;
  sg.blk_case = bt.blk_case;
  sg.active = bt.active;
  return sg;
}

// Item ranges of a compacted launch: virtual workgroup vb covers items
// [blk[vb], blk[G + vb]) (at most blockDim items), and only the ranges of
// cases still iterating are listed, so a launch over a large, sparsely used
// arena costs what its active cases hold.  blk == NULL: one item per thread
// over the whole index range.
struct Items {
  const int32_t* blk;  // [2][G] starts, then ends
  int G;
};

__device__ inline int item_index(const Items& it) {
  if (!it.blk) return xcd_item();
  const int vb = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int i = it.blk[vb] + (int)threadIdx.x;
  return i < it.blk[it.G + vb] ? i : 0x7fffffff;   // past the range: callers' bound check fails
}

// Workgroups of a launch over n items (0: nothing to launch).
inline unsigned items_grid(const Items& it, int64_t n) {
  if (it.blk) return (unsigned)it.G;
  return n > 0 ? grid_for(n) : 0u;
}

// Per-cell / per-face case lookups of a batch for the face and cell kernels.
// With cell_case == NULL there is one case, always active, and the scalar
// kernel arguments (viscosity, reference row) apply.
struct BatchDev {
  const int32_t* cell_case;  // [C] case of each cell, -1 for unused rows
  const int32_t* face_case;  // [F]
  const uint8_t* active;     // [K]
  const real* nu;           // [K] or NULL
  const int32_t* ref_cell;   // [K] or NULL
  const real* ref_value;    // [K]
  Items cells, faces, bfaces;  // compacted launches (cells, all faces, boundary faces)
  __device__ bool cell_skip(int c) const {
    if (!cell_case) return false;
    const int k = cell_case[c];
    return k < 0 || !active[k];
  }
  __device__ bool face_skip(int f) const {
    if (!face_case) return false;
    const int k = face_case[f];
    return k < 0 || !active[k];
  }
};

inline BatchDev batch_dev(const wcfd_batch* bt) {
  BatchDev d{};
  if (bt) {
    d.cell_case = bt->cell_case;
    d.face_case = bt->face_case;
    d.active = bt->active;
    d.nu = bt->nu;
    d.ref_cell = bt->ref_cell;
    d.ref_value = bt->ref_value;
    d.cells = Items{bt->cell_blk, bt->n_cell_blk};
    d.faces = Items{bt->face_blk, bt->n_face_blk};
    d.bfaces = Items{bt->bface_blk, bt->n_bface_blk};
  }
  return d;
}

// Thread t's share of a fold: s = ((0 + p[t]) + p[t+B]) + p[t+2B] ...,
// with the loads of U terms issued before their additions (one memory round
// trip per U terms instead of per term); the order of the additions, and so
// the sum, does not depend on U.
template <int ND, int U>
__device__ __forceinline__ void fold_thread(const double* const (&p)[ND], int nb,
                                            double (&s)[ND]) {
  const int step = blockDim.x;
  int b = threadIdx.x;
  for (; b + (U - 1) * step < nb; b += U * step) {
    double v[U][ND];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int i = 0; i < ND; ++i) v[u][i] = p[i][b + u * step];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int i = 0; i < ND; ++i) s[i] += v[u][i];
  }
  for (; b < nb; b += step) {
#pragma unroll
    for (int i = 0; i < ND; ++i) s[i] += p[i][b];
  }
}

// Loads in flight per thread in the folds inside iteration kernels, where
// they add to the registers held by the prefetched row (build-time choice;
// the prefold kernel uses 16).
#ifndef WCFD_FOLD_UNROLL
#define WCFD_FOLD_UNROLL 2
#endif
constexpr int kFoldUnroll = WCFD_FOLD_UNROLL;
// the same in the prefold kernel, which runs alone (one workgroup per case):
// 4, 8 and 16 measured within 4 % of each other at 4.4 and 8.6 M cells,
// where the kernel is bound by its one CU's read rate; 24 and 32 were slower
#ifndef WCFD_PREFOLD_UNROLL
#define WCFD_PREFOLD_UNROLL 16
#endif

// Sum of nb partials (the same order in every workgroup); valid in every
// thread.
template <int U = kFoldUnroll>
__device__ inline double fold(const double* __restrict__ part, int nb) {
  __shared__ double lds[kMaxWaves];
  __shared__ double res;
  const double* const pp[1] = {part};
  double sv[1] = {0.0};
  fold_thread<1, U>(pp, nb, sv);
  double s = wave_sum(sv[0]);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += lds[w];
    res = t;
  }
  __syncthreads();
  return res;   // a following fold writes res only after its own barrier
}

// ND folds in one pass: out[i] = sum of p[i][0 .. nb) (the order of fold());
// one pair of barriers for all of them.
template <int ND, int U = kFoldUnroll>
__device__ inline void fold_n(const double* const (&p)[ND], int nb, double (&out)[ND]) {
  __shared__ double lds[ND * kMaxWaves];
  __shared__ double res[ND];
  double s[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) s[i] = 0.0;
  fold_thread<ND, U>(p, nb, s);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  wave_sum_n<ND>(s);
#pragma unroll
  for (int i = 0; i < ND; ++i)
    if (lane == 0) lds[i * kMaxWaves + wid] = s[i];
  __syncthreads();
  if (threadIdx.x < ND) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += lds[threadIdx.x * kMaxWaves + w];
    res[threadIdx.x] = t;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ND; ++i) out[i] = res[i];
}

// A fold_n whose loads start early.  Inside an iteration kernel the fold of
// the previous kernel's partials and the row's own gathers are independent,
// but vector loads retire in issue order: partials loaded after the row's
// gathers can only be summed once every gather has landed.  load() issues a
// thread's first E terms per sum (all of them when nb <= E * blockDim) before
// the caller issues the row's loads; finish() adds them, and any further
// terms, in fold_thread's order and reduces the block as fold_n does, so the
// result is bitwise fold_n's.  The wave and block reduction then overlaps the
// gathers still in flight.
template <int ND, int E = kFoldUnroll>
struct EarlyFold {
  double v[E][ND];
  __device__ __forceinline__ void load(const double* const (&p)[ND], int nb) {
    const int step = blockDim.x;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int b = (int)threadIdx.x + u * step;
#pragma unroll
      for (int i = 0; i < ND; ++i) v[u][i] = b < nb ? p[i][b] : 0.0;
    }
  }
  __device__ __forceinline__ void finish(const double* const (&p)[ND], int nb,
                                         double (&out)[ND]) {
    __shared__ double lds[ND * kMaxWaves];
    __shared__ double res[ND];
    const int step = blockDim.x;
    double s[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) s[i] = 0.0;
#pragma unroll
    for (int u = 0; u < E; ++u)
      if ((int)threadIdx.x + u * step < nb) {
#pragma unroll
        for (int i = 0; i < ND; ++i) s[i] += v[u][i];
      }
    for (int b = (int)threadIdx.x + E * step; b < nb; b += step) {
#pragma unroll
      for (int i = 0; i < ND; ++i) s[i] += p[i][b];
    }
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    wave_sum_n<ND>(s);
#pragma unroll
    for (int i = 0; i < ND; ++i)
      if (lane == 0) lds[i * kMaxWaves + wid] = s[i];
    __syncthreads();
    if (threadIdx.x < ND) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += lds[threadIdx.x * kMaxWaves + w];
      res[threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ND; ++i) out[i] = res[i];
  }
};

// Block sum of ND values; thread 0 stores them as partials part[i*G + vb]
// (COH: at agent scope, for readers on other XCDs inside the same launch).
template <int ND, bool COH = false>
__device__ inline void store_partials(double (&v)[ND], double* __restrict__ part, int G, int vb) {
  __shared__ double lds[ND * kMaxWaves];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  wave_sum_n<ND>(v);
#pragma unroll
  for (int i = 0; i < ND; ++i)
    if (lane == 0) lds[i * kMaxWaves + wid] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += lds[i * kMaxWaves + w];
      if constexpr (COH)
        __hip_atomic_store(part + i * G + vb, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      else
        part[i * G + vb] = t;
    }
  }
}

// ------------------------------------------------------- matrix access
// Krylov kernels are written once against these two row views.
//
// CsrMat: a general CSR matrix (row_ptr, col, values, diagonal position).
struct CsrMat {
  static constexpr bool kEll = false;
  static constexpr int kRow = 7;      // register arrays of the prefetching forms (unused)
  int n;
  const int32_t* rp;
  const int32_t* col;
  const real* val;
  const int32_t* dpos;
  __device__ real diag(int i) const { return val[dpos[i]]; }
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
// band fits), halving the index bytes.  NS slots are read (6 hexahedra, 5
// wedges, 4 tetrahedra); the planes of slots NS..5 hold zeros and are
// skipped.
template <bool D16, int NS = 6>
struct EllMat {
  static constexpr bool kEll = true;
  static constexpr int kSlots = NS;
  static constexpr int kRow = NS + 1;   // diagonal + slots
  int n;
  const real* val;
  const int32_t* nbr;
  const int16_t* dnb;
  __device__ real diag(int i) const { return val[i]; }
  // the row as 7 (coefficient, column) pairs in row() order, for kernels
  // that issue every load of a row before they wait on anything
  __device__ void load(int i, real (&a)[kRow], int (&c)[kRow]) const {
    a[0] = val[i];
    c[0] = i;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      c[k + 1] = D16 ? i + (int)dnb[k * n + i] : nbr[k * n + i];
      a[k + 1] = val[(k + 1) * n + i];
    }
  }
  // The same row with the columns as byte offsets of real vectors
  // (o[j] = 4 c[j]); every load is a uniform base plus a 32-bit offset.
  __device__ void load_off(int i, real (&a)[kRow], uint32_t (&o)[kRow]) const {
    const uint32_t oi = (uint32_t)i * 4u;
    a[0] = ld_mat(val, oi);
    o[0] = oi;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const size_t col0 = (size_t)k * (size_t)n;
      const int cj = D16 ? i + (int)ld_mat(dnb + col0, (uint32_t)i * 2u)
                         : ld_mat(nbr + col0, oi);
      o[k + 1] = (uint32_t)cj * 4u;
      a[k + 1] = ld_mat(val + col0 + (size_t)n, oi);
    }
  }
  template <class F>
  __device__ void row(int i, F&& f) const {
    f(val[i], i);
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int c = D16 ? i + (int)dnb[k * n + i] : nbr[k * n + i];
      f(val[(k + 1) * n + i], c);
    }
  }
};

}  // namespace wcfd
