// MFMA GEMM for gfx950 (CDNA4):  D(m, n) = sum_k A(m, k) B(k, n)
//
// Both operands are read along k: A is row-major (A(m, k) = A[m * lda + k],
// i.e. a column-major matrix used transposed) and B is column-major
// (B(k, n) = B[k + n * ldb]). The Gibbs products of the batched solvers are of
// that form once the kernel is kept in both orientations: K v uses K^T stored
// column-major as the row-major K, and K^T u uses K stored column-major.
//
// Workgroup tile BM x BN over k-steps of BK, staged through LDS in k-major
// rows (each row of BK contiguous k values, padded against bank conflicts);
// 4 wavefronts in a 2 x 2 grid, each computing a (BM/2) x (BN/2) block with
// 16 x 16 MFMA tiles. The next k-step is loaded from global memory while the
// current one is multiplied and stored into the second LDS stage after it. D is written through an epilogue functor
// op(m, n, value) so callers can fuse their elementwise update.
#pragma once

#include <algorithm>
#include <type_traits>
#include <utility>

#include "common.hpp"

namespace rwig {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef double doublex4_t __attribute__((ext_vector_type(4)));

// One MFMA flavour: element type, k per instruction, fragment of one lane.
//   bf16: v_mfma_f32_16x16x32_bf16, A/B fragments of 8 consecutive k
//   fp32: v_mfma_f32_16x16x4_f32,   A/B fragments of 1 k
//   fp64: v_mfma_f64_16x16x4_f64,   A/B fragments of 1 k
// Lane l of a 16 x 16 tile: A row l % 16, B column l % 16, k block l / 16;
// accumulator register r < 4 of lane l: column l % 16 and row
//   bf16, fp32: 4 (l / 16) + r        fp64: 4 r + l / 16
// (the f64 layout is the one composable_kernel's xdlops tables give:
// group size 1, four input blocks).
template <typename T> struct Mfma;

template <> struct Mfma<uint16_t> {
  using elem = uint16_t;
  using frag_t = bf16x8_t;
  using acc_t = floatx4_t;
  using out_t = float;
  static constexpr int KI = 32;  // k per instruction
  static constexpr int KL = 8;   // k per lane
  __device__ static acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static frag_t load(const elem *p) { return *reinterpret_cast<const frag_t *>(p); }
  __device__ static void mma(acc_t &c, const frag_t &a, const frag_t &b) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <> struct Mfma<double> {
  using elem = double;
  using frag_t = double;
  using acc_t = doublex4_t;
  using out_t = double;
  static constexpr int KI = 4;
  static constexpr int KL = 1;
  __device__ static acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  __device__ static frag_t load(const elem *p) { return *p; }
  __device__ static void mma(acc_t &c, const frag_t &a, const frag_t &b) {
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return 4 * r + (lane >> 4); }
};

template <> struct Mfma<float> {
  using elem = float;
  using frag_t = float;
  using acc_t = floatx4_t;
  using out_t = float;
  static constexpr int KI = 4;
  static constexpr int KL = 1;
  __device__ static acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static frag_t load(const elem *p) { return *p; }
  __device__ static void mma(acc_t &c, const frag_t &a, const frag_t &b) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Tile shape of a product: workgroup tile BM x BN over k-steps of BK, LDS rows
// padded by PAD elements, WGM x WGN wavefronts each computing a
// (BM/WGM) x (BN/WGN) block of 16 x 16 MFMA tiles, PF k-steps of global loads
// in flight (PF = 2: a second register set, so that a load has two k-steps of
// MFMA work to arrive in).
// DMA = 1: the tiles go from global memory straight into LDS
// (global_load_lds_dwordx4, no register staging) through PF + 1 LDS stages, PF
// k-steps ahead, in an unpadded layout whose 16-byte chunks are XOR-swizzled
// by row (PAD is then unused). Needs tile rows of 128 or 256 bytes (BF16: 64
// or 128 k; FP32: 32 or 64 k; FP64: 16 or 32 k) and 16-byte aligned rows;
// gemm_tn falls back to the register path otherwise.
template <typename T, int BM_, int BN_, int BK_, int PAD_, int WGM_, int WGN_, int PF_ = 1,
          int DMA_ = 0>
struct GemmShape {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, PAD = PAD_, WGM = WGM_, WGN = WGN_;
  static constexpr int PF = PF_;  // k-steps of global loads in flight (>= 1)
  static constexpr int DMA = DMA_;
  static constexpr int NT = 64 * WGM * WGN;  // threads per workgroup
  static constexpr int WM = BM / WGM, WN = BN / WGN;
  static constexpr int TM = WM / 16, TN = WN / 16;
  static constexpr int LDL = BK + PAD;
  static_assert(WM % 16 == 0 && WN % 16 == 0, "wave block must be a multiple of 16");
  static_assert(BK % Mfma<T>::KI == 0, "k-step must be a multiple of the MFMA k");
  static_assert(!DMA || ((BK * sizeof(T) == 128 || BK * sizeof(T) == 256) &&
                         BM % (WGM * WGN * 1024 / (BK * sizeof(T))) == 0 &&
                         BN % (WGM * WGN * 1024 / (BK * sizeof(T))) == 0),
                "the direct-to-LDS path takes tile rows of 128 or 256 bytes (8 or 16 chunks "
                "of 16 bytes) that split into 1 KB moves per wavefront");
};

template <typename T> struct GemmCfg;
template <> struct GemmCfg<float> : GemmShape<float, 128, 128, 16, 4, 2, 2> {};
// BF16: 8 wavefronts (2 x 4, 64 x 32 each), 3 k-steps of direct-to-LDS
// moves in flight (4 LDS stages, 128 KB). 16 wavefronts of 32 x 32 read a
// third more LDS bytes per flop (128 against 96 KB of fragments per k-step)
// and ran 35.0 against 33.3 us at 4096 x 1024 x 4096 with the same loop.
template <> struct GemmCfg<uint16_t> : GemmShape<uint16_t, 128, 128, 64, 16, 2, 4, 3, 1> {};
template <> struct GemmCfg<double> : GemmShape<double, 64, 64, 16, 2, 2, 2> {};

// FP64 products with many columns: 128 x 128 tiles, 8 wavefronts (2 x 4),
// direct-to-LDS moves three k-steps ahead (72.7 TFLOP/s on 4096 x 1024 x 4096,
// 0.92 of the FP64 MFMA peak, against 64.7 with 16 wavefronts and 55 for the
// 64 x 64 register tiles). Each output element sees the same MFMA sequence as
// with GemmCfg<double>, so results are identical.
using GemmF64Wide = GemmShape<double, 128, 128, 16, 2, 2, 4, 3, 1>;
// FP32 products with many columns (the WDL step's 2000 x 8192 x 2000):
// 128 x 128 tiles, 8 wavefronts, direct-to-LDS moves two k-steps ahead, 32 k
// per step (482.7 us against 534.9 us for the 4-wavefront register tiles of
// GemmCfg<float>, same MFMA sequence per output element).
using GemmF32Wide = GemmShape<float, 128, 128, 32, 4, 2, 4, 2, 1>;
// the tiles of a product over T with many columns
template <typename T> struct WideCfg { using type = GemmCfg<T>; };
template <> struct WideCfg<double> { using type = GemmF64Wide; };
template <> struct WideCfg<float> { using type = GemmF32Wide; };

// Loads the BM x BK (or BN x BK) tile rows starting at (r0, k0) of a matrix
// stored with its k values contiguous (element (r, k) at p[r * ld + k]) into
// a register set; rows >= nrows and k >= K read as zero. Vector width: 16
// bytes. The row pointers are shared by all the register sets of a product
// (a set per k-step of loads in flight).
template <typename T, int ROWS, int BK, int NT>
struct TileLoader {
  static constexpr int VEC = 16 / sizeof(T);
  static constexpr int CHUNKS = ROWS * BK / VEC;      // 16-byte chunks per tile
  static constexpr int PER_THREAD = (CHUNKS + NT - 1) / NT;
  static constexpr bool EVEN = CHUNKS % NT == 0;      // every thread has PER_THREAD chunks
  using Regs = uint4[PER_THREAD];
  const T *row_ptr[PER_THREAD];  // start of this thread's rows (a valid row if outside)
  unsigned inside = 0;           // bit c: chunk c lies on a row of the matrix

  __device__ static bool has(int idx) { return EVEN || idx < CHUNKS; }

  // rows r0.. of a matrix with contiguous k (element (r, k) at p[r * ld + k]);
  // rmap (optional): tile row r reads stored row rmap[r]. Resolved once.
  // Chunks outside the matrix read its first row and are zeroed, so that the
  // loads of a k-step carry no branch (a branch around them makes the
  // compiler wait for each load at once instead of after the MFMA work).
  __device__ void init(const T *__restrict__ p, long ld, int nrows, int r0,
                       const int *__restrict__ rmap) {
#pragma unroll
    for (int c = 0; c < PER_THREAD; ++c) {
      const int idx = threadIdx.x + c * NT;
      const int r = r0 + idx / (BK / VEC);
      const bool in = has(idx) && r < nrows;
      if (in) inside |= 1u << c;
      row_ptr[c] = p + (in ? (rmap ? (long)rmap[r] : (long)r) * ld : 0L);
    }
  }
  // loads k values k0 .. k0 + BK of the tile's rows; needs k0 + BK <= K
  __device__ void load_full(Regs &reg, int k0) const {
#pragma unroll
    for (int c = 0; c < PER_THREAD; ++c) {
      const int idx = threadIdx.x + c * NT;
      if (!has(idx)) continue;
      const int k = k0 + (idx % (BK / VEC)) * VEC;
      reg[c] = *reinterpret_cast<const uint4 *>(row_ptr[c] + k);  // zeroed in store()
    }
  }
  // the last, partial k-step: k >= K read as zero
  __device__ void load_tail(Regs &reg, int K, int k0) const {
#pragma unroll
    for (int c = 0; c < PER_THREAD; ++c) {
      const int idx = threadIdx.x + c * NT;
      const int k = k0 + (idx % (BK / VEC)) * VEC;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (has(idx) && ((inside >> c) & 1u)) {
        const T *src = row_ptr[c] + k;
        if (k + VEC <= K) {
          v = *reinterpret_cast<const uint4 *>(src);
        } else {
          T tmp[VEC];
#pragma unroll
          for (int e = 0; e < VEC; ++e) tmp[e] = (k + e < K) ? src[e] : T(0);
          v = *reinterpret_cast<const uint4 *>(tmp);
        }
      }
      reg[c] = v;
    }
  }
  // the loaded chunks into the LDS tile; chunks outside the matrix as zero
  // (applied here, after the k-step's MFMA work, not at the load)
  __device__ void store(const Regs &reg, T *lds, int ldl) const {
#pragma unroll
    for (int c = 0; c < PER_THREAD; ++c) {
      const int idx = threadIdx.x + c * NT;
      if (has(idx)) {
        const int row = idx / (BK / VEC), kc = (idx % (BK / VEC)) * VEC;
        const bool in = (inside >> c) & 1u;
        const uint4 v = reg[c];
        *reinterpret_cast<uint4 *>(lds + row * ldl + kc) =
            make_uint4(in ? v.x : 0u, in ? v.y : 0u, in ? v.z : 0u, in ? v.w : 0u);
      }
    }
  }
};

// Deferred epilogue, with an optional split of the k range over blockIdx.z
// (skinny products: few row tiles, long k). With a workspace, each split
// writes its partial tile there and a second kernel adds the partials in
// split order and runs the epilogue with more workgroups than the product
// has (the epilogue's loads are latency-bound at one workgroup per tile).
// Without one (ks = 1), the epilogue runs in the product's kernel.
template <typename T> struct SplitK {
  int ks = 1;
  typename Mfma<T>::out_t *work = nullptr;  // ks x tiles x (per-thread tile) x 256
};

// k splits for an M x K operand: enough workgroups along M x splits to cover
// the chip when M alone gives few tiles, at least 4 k-steps per split. A
// function of (M, K) only, so that the arithmetic of a column does not depend
// on how many columns the product has.
template <typename T, class CF = GemmCfg<T>> inline int splitk_for(int M, int K) {
  const int tiles = (M + CF::BM - 1) / CF::BM;
  const int steps = (K + CF::BK - 1) / CF::BK;
  int ks = 256 / tiles;
  ks = std::min(ks, 8);
  ks = std::min(ks, steps / 4);
  return std::max(ks, 1);
}

// Workspace elements of a split product with M rows, up to N columns and ks
// splits.
template <typename T, class CF = GemmCfg<T>>
inline size_t splitk_work_elems(int M, int N, int ks) {
  const size_t tiles = (size_t)((M + CF::BM - 1) / CF::BM) * ((N + CF::BN - 1) / CF::BN);
  return (size_t)ks * tiles * CF::BM * CF::BN;
}

// Tiles of a product by its number of columns, from sweeps at
// 4096 x cols x 4096 on MI355X: fewer columns take smaller tiles, so that the
// product still spreads over the CUs. Every output element sees the same
// sequence of MFMAs in every tile shape (for a given k split), so a column's
// arithmetic does not depend on the tiles it falls in.
//   BF16 (us, with the deferred epilogue)      FP64 (us)
//   cols <=  32: 16 x 32, 2 waves     15.3-15.5
//        <= 128: 32 x 32, 4 waves     16-20     <= 256: 32 x 32, 4 waves   100-210
//        <= 256: 32 x 32, 1 wave      23-25     <= 512: 128 x 64, 8 waves  310-320
//        <= 512: 64 x 64, 4 waves     31-34     <= 768: 64 x 64, 4 waves   495-510
//        above: 128 x 128, 16 waves   44-50     above: 128 x 128, 16 waves 560-590
// (the 128 x 128 tiles alone take 37-42 us in BF16 and 530-555 us in FP64 at
// any count up to 512 columns). FP32 products keep GemmCfg<float> here; the
// WDL step takes GemmF32Wide.
using Bf16Micro = GemmShape<uint16_t, 16, 32, 64, 16, 1, 2, 3, 1>;
using Bf16Tiny = GemmShape<uint16_t, 32, 32, 64, 16, 2, 2, 3, 1>;
using Bf16Small = GemmShape<uint16_t, 32, 32, 64, 16, 1, 1, 2, 1>;
using Bf16Mid = GemmShape<uint16_t, 64, 64, 64, 16, 2, 2, 2, 1>;
using F64Small = GemmShape<double, 32, 32, 16, 2, 2, 2, 3, 1>;
using F64Mid = GemmShape<double, 128, 64, 16, 2, 4, 2, 3, 1>;

template <typename T, class F> void with_tiles(int cols, F f) {
  if constexpr (std::is_same<T, uint16_t>::value) {
    if (cols <= 32) f(Bf16Micro{});
    else if (cols <= 128) f(Bf16Tiny{});
    else if (cols <= 256) f(Bf16Small{});
    else if (cols <= 512) f(Bf16Mid{});
    else f(GemmCfg<uint16_t>{});
  } else if constexpr (std::is_same<T, double>::value) {
    if (cols <= 256) f(F64Small{});
    else if (cols <= 512) f(F64Mid{});
    else if (cols <= 768) f(GemmCfg<double>{});
    else f(GemmF64Wide{});
  } else {
    f(GemmCfg<T>{});
  }
}

// split workspace elements of a rows x (up to cols) product in any of the
// tile shapes with_tiles picks
template <typename T> inline size_t tiles_work_elems(int rows, int cols, int ks) {
  size_t n = 0;
  for (int c : {1, 33, 129, 257, 513, 769})
    with_tiles<T>(c, [&](auto cfg) {
      n = std::max(n, splitk_work_elems<T, decltype(cfg)>(rows, cols, ks));
    });
  return n;
}

// Epilogue of one workgroup tile: op(m, n or nmap[n], value) for the tile's
// elements inside M x N; accumulator lane layout: column lane % 16, row
// Mfma::row(lane, r).
template <typename T, class CF, class Op>
__device__ __forceinline__ void gemm_epilogue(
    int M, int N, int m0, int n0, const int *__restrict__ nmap,
    const typename Mfma<T>::acc_t (&acc)[CF::TM][CF::TN], Op &op) {
  using MF = Mfma<T>;
  constexpr int WM = CF::WM, WN = CF::WN, TM = CF::TM, TN = CF::TN;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wm = wv % CF::WGM, wn = wv / CF::WGM, lr = lane & 15;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn * WN + j * 16 + lr;
      if (n >= N) continue;
      const int nn = nmap ? nmap[n] : n;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * WM + i * 16 + MF::row(lane, r);
        if (m < M) op(m, nn, acc[i][j][r]);
      }
    }
}

#define RWIG_LDS __attribute__((address_space(3)))

// 16 bytes per lane from global memory into LDS at l + 16 * lane (l uniform
// across the wavefront): global_load_lds_dwordx4. A plain device function, so
// that the host pass of a template that calls it never sees the builtin.
__device__ __forceinline__ void lds_move16(const void *g, RWIG_LDS void *l) {
  __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
}

// s_waitcnt vmcnt(v) with the other counters left alone (gfx9 encoding: vmcnt
// in bits 3:0 and 15:14, expcnt 7 in 6:4, lgkmcnt 15 in 11:8)
template <int V> __device__ __forceinline__ void wait_vmcnt() {
  static_assert(V >= 0 && V < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((V & 15) | (7 << 4) | (15 << 8) | ((V >> 4) << 14));
}

// Waits until at most `later` k-steps of MOVES moves each are outstanding
// (later <= L, a compile-time count per branch for s_waitcnt).
template <int MOVES, int L> __device__ __forceinline__ void wait_later(int later) {
  if constexpr (L == 0) {
    wait_vmcnt<0>();
  } else {
    if (later >= L) wait_vmcnt<L * MOVES>();
    else wait_later<MOVES, L - 1>(later);
  }
}

// Main loop, register path: global loads into PF register sets, stored into
// one of two padded LDS stages after the MFMA work of the current k-step.
template <typename T, class CF>
__device__ __forceinline__ void gemm_main_regs(
    int M, int N, int K, const T *__restrict__ A, long lda, const T *__restrict__ B, long ldb,
    const int *__restrict__ nmap, int m0, int n0, int kbeg, int kend,
    typename Mfma<T>::acc_t (&acc)[CF::TM][CF::TN]) {
  using MF = Mfma<T>;
  constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, LDL = CF::LDL;
  constexpr int WM = CF::WM, WN = CF::WN;  // per-wave block
  constexpr int TM = CF::TM, TN = CF::TN;
  // two LDS stages: the next k-step is stored into one while the other is
  // being multiplied, one barrier per k-step
  __shared__ __attribute__((aligned(16))) T As[2][BM * LDL];
  __shared__ __attribute__((aligned(16))) T Bs[2][BN * LDL];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wm = wv % CF::WGM, wn = wv / CF::WGM;
  using LA = TileLoader<T, BM, BK, CF::NT>;
  using LB = TileLoader<T, BN, BK, CF::NT>;
  LA la;
  LB lb;
  la.init(A, lda, M, m0, nullptr);
  lb.init(B, ldb, N, n0, nmap);
  const int lr = lane & 15, lk = (lane >> 4) * MF::KL;
  // one k-step of MFMA work on an LDS stage
  auto mma_step = [&](const T *as, const T *bs) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += MF::KI) {
      typename MF::frag_t af[TM], bfr[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = MF::load(as + (wm * WM + i * 16 + lr) * LDL + kk + lk);
#pragma unroll
      for (int j = 0; j < TN; ++j) bfr[j] = MF::load(bs + (wn * WN + j * 16 + lr) * LDL + kk + lk);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) MF::mma(acc[i][j], af[i], bfr[j]);
    }
  };
  // global loads of the k-step at kk0 into a register set: one uniform branch
  // around both loaders, so that the full-step loads form one block with no
  // join between them and the compiler defers their waits to the LDS stores
  // that follow the MFMA work
  auto issue = [&](typename LA::Regs &ra, typename LB::Regs &rb, int kk0) {
    if (kk0 + BK <= K) {
      la.load_full(ra, kk0);
      lb.load_full(rb, kk0);
    } else {
      la.load_tail(ra, K, kk0);
      lb.load_tail(rb, K, kk0);
    }
  };
  // PF register sets: set q holds k-step s with s % PF == q. At step s the
  // LDS stage s % 2 holds step s, sets (s+1 .. s+PF-1) % PF have their loads
  // in flight, and set s % PF (stored to LDS at step s - 1) takes step s + PF;
  // after the MFMA work, step s + 1 goes from its set to the other stage. One
  // barrier per k-step.
  constexpr int PF = CF::PF;
  typename LA::Regs ra[PF];
  typename LB::Regs rb[PF];
  const int nst = (kend - kbeg + BK - 1) / BK;
#pragma unroll
  for (int q = 0; q < PF; ++q)
    if (q < nst) issue(ra[q], rb[q], kbeg + q * BK);
  if (nst > 0) {
    la.store(ra[0], As[0], LDL);
    lb.store(rb[0], Bs[0], LDL);
  }
  __syncthreads();
  for (int s0 = 0; s0 < nst; s0 += PF) {
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int s = s0 + q;
      if (s < nst) {
        if (s + PF < nst) issue(ra[q], rb[q], kbeg + (s + PF) * BK);
        mma_step(As[s & 1], Bs[s & 1]);
        if (s + 1 < nst) {
          la.store(ra[(q + 1) % PF], As[(s + 1) & 1], LDL);
          lb.store(rb[(q + 1) % PF], Bs[(s + 1) & 1], LDL);
        }
        __syncthreads();
      }
    }
  }
}

// s_waitcnt lgkmcnt(0) with the other counters left alone
__device__ __forceinline__ void wait_lgkmcnt0() {
  __builtin_amdgcn_s_waitcnt(15 | (7 << 4) | (0 << 8) | (3 << 14));
}

// f(integral_constant<int, I>) for I = 0 .. sizeof...(I) - 1, in order
template <int... I, class F>
__device__ __forceinline__ void unroll_seq(std::integer_sequence<int, I...>, F &&f) {
  (f(std::integral_constant<int, I>{}), ...);
}

// Main loop, direct-to-LDS path (CF::DMA): each global_load_lds_dwordx4
// moves 1 KB (64 lanes x 16 bytes, written to LDS in lane order): 8 tile rows
// of 64 k or 4 of 128 k. The lane with slot s of tile row r loads the 16-byte
// k chunk s ^ (r % 8), so that chunk c of row r sits at slot c ^ (r % 8) and
// the MFMA fragment reads (16 rows, one chunk) spread over all LDS banks.
// Rows past the matrix read a valid row instead: they only feed outputs the
// epilogue discards.
//
// PF + 1 LDS stages and two register sets of fragments. Step t: wait for this
// wavefront's moves of step t + 1 (those of later steps stay in flight) and
// for its fragment reads of step t, one barrier, then the moves of step
// t + PF + 1 into the stage of step t (every wavefront holds step t's
// fragments in registers by then), the fragment reads of step t + 1 into the
// other set, and the MFMAs of step t. Each output element sees the MFMAs of
// the register path in the same order.
//
// Away from the end of the k range the loop is unrolled over a whole cycle of
// stages, so that every stage offset is a constant (the fragment reads take
// immediate LDS offsets and no address arithmetic), and the step's
// instructions are issued interleaved: one MFMA, then one move, for each of
// the wavefront's moves; then one MFMA, then one fragment read; then the rest.
// A wavefront then keeps the matrix core fed while its moves and reads go
// out, which matters most with one or two wavefronts per SIMD (BF16 4096 x
// 1024 x 4096 on 128 x 128 tiles: 37.8 -> 33.3 us with 8 wavefronts instead
// of 16; FP64: 531 -> 472 us).
template <typename T, class CF, class Acc>
__device__ __forceinline__ void gemm_main_dma(
    int M, int N, const T *__restrict__ A, long lda, const T *__restrict__ B, long ldb,
    const int *__restrict__ nmap, int m0, int n0, int kbeg, int kend, Acc &acc) {
  using MF = Mfma<T>;
  using frag_t = typename MF::frag_t;
  constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK;
  constexpr int WM = CF::WM, WN = CF::WN, TM = CF::TM, TN = CF::TN;
  constexpr int NW = CF::WGM * CF::WGN;
  constexpr int ROWB = BK * (int)sizeof(T);  // bytes of a tile row: 8 or 16 chunks
  constexpr int CPR = ROWB / 16;             // chunks per row
  constexpr int EPC = 16 / (int)sizeof(T);   // elements per chunk
  constexpr int RPM = 64 / CPR;              // rows per move (64 lanes x 16 bytes)
  constexpr int PF = CF::PF, STAGES = PF + 1;
  constexpr int STAGEB = (BM + BN) * ROWB;
  constexpr int MA = BM / NW / RPM, MB = BN / NW / RPM;  // moves per wavefront and k-step
  constexpr int MOVES = MA + MB;
  constexpr int KS = BK / MF::KI;  // MFMA k-slices per step
  constexpr int NM = KS * TM * TN, NR = KS * (TM + TN);  // MFMAs and fragment reads per step
  static_assert(PF * MOVES < 64, "the moves in flight must fit vmcnt");
  __shared__ __attribute__((aligned(1024))) char smem[STAGES * STAGEB];
  RWIG_LDS char *lds = (RWIG_LDS char *)smem;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wv % CF::WGM, wn = wv / CF::WGM;
  // the lane's row within the move and its LDS slot; the chunk it loads is
  // slot ^ (tile row % 8)
  const int sub = lane / CPR, slot = lane % CPR;
  const T *pa[MA], *pb[MB];
#pragma unroll
  for (int q = 0; q < MA; ++q) {
    const int rl = (wv * MA + q) * RPM + sub, r = m0 + rl;
    pa[q] = A + (long)(r < M ? r : 0) * lda + (slot ^ (rl & 7)) * EPC;
  }
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int rl = (wv * MB + q) * RPM + sub, n = n0 + rl;
    const long r = n < N ? (nmap ? nmap[n] : n) : (nmap ? nmap[0] : 0);
    pb[q] = B + r * ldb + (slot ^ (rl & 7)) * EPC;
  }
  const int nst = (kend - kbeg) / BK;  // whole k-steps; a shorter last one follows
  auto issue = [&](int stage, int s) {
    RWIG_LDS char *st = lds + stage * STAGEB;
    const int k = kbeg + s * BK;
#pragma unroll
    for (int q = 0; q < MA; ++q)
      lds_move16(pa[q] + k, (RWIG_LDS void *)(st + (wv * MA + q) * RPM * ROWB));
#pragma unroll
    for (int q = 0; q < MB; ++q)
      lds_move16(pb[q] + k, (RWIG_LDS void *)(st + (BM + (wv * MB + q) * RPM) * ROWB));
  };
  // byte offset within a stage of the lane's fragment of tile row `row` at
  // k-slice ks: chunk k / EPC (at its swizzled slot), element k % EPC in it
  const int lr = lane & 15, lk = (lane >> 4) * MF::KL;  // the lane's row and first k
  auto foff = [&](int row, int ks) {
    const int k = ks * MF::KI + lk;
    return row * ROWB + (((k / EPC) ^ (row & 7)) << 4) + (k % EPC) * (int)sizeof(T);
  };
  int oa[KS][TM], ob[KS][TN];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
    for (int i = 0; i < TM; ++i) oa[ks][i] = foff(wm * WM + i * 16 + lr, ks);
#pragma unroll
    for (int j = 0; j < TN; ++j) ob[ks][j] = foff(BM + wn * WN + j * 16 + lr, ks);
  }
  frag_t fa[2][KS][TM], fb[2][KS][TN];
  auto read = [&](auto set_c, int stage) {
    constexpr int SET = decltype(set_c)::value;
    const RWIG_LDS char *st = lds + stage * STAGEB;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
        fa[SET][ks][i] = *reinterpret_cast<const RWIG_LDS frag_t *>(st + oa[ks][i]);
#pragma unroll
      for (int j = 0; j < TN; ++j)
        fb[SET][ks][j] = *reinterpret_cast<const RWIG_LDS frag_t *>(st + ob[ks][j]);
    }
  };
  auto mma = [&](auto set_c) {
    constexpr int SET = decltype(set_c)::value;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) MF::mma(acc[i][j], fa[SET][ks][i], fb[SET][ks][j]);
  };
  auto barrier = [] {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };
  // The last, partial k-step (k range not a multiple of BK): the same lanes
  // load the same 16-byte chunks through registers, k >= kend as zero, and
  // store them where the moves would have put them; the MFMAs then run as in
  // any other step, so each output element still sees the register path's
  // MFMA sequence (the zero k-slices of the step add exact zeros).
  auto tail = [&](int k0) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    auto chunk = [&](const T *row, int kc) {
      u32x4 v = {0u, 0u, 0u, 0u};
      if (kc + EPC <= kend) {
        v = *reinterpret_cast<const u32x4 *>(row);
      } else if (kc < kend) {
        T tmp[EPC];
#pragma unroll
        for (int e = 0; e < EPC; ++e) tmp[e] = kc + e < kend ? row[e] : T(0);
        v = *reinterpret_cast<const u32x4 *>(tmp);
      }
      return v;
    };
    wait_vmcnt<0>();
    wait_lgkmcnt0();
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MA; ++q) {
      const int rl = (wv * MA + q) * RPM + sub;
      const u32x4 v = chunk(pa[q] + k0, k0 + (slot ^ (rl & 7)) * EPC);
      *reinterpret_cast<RWIG_LDS u32x4 *>(lds + (wv * MA + q) * RPM * ROWB + lane * 16) = v;
    }
#pragma unroll
    for (int q = 0; q < MB; ++q) {
      const int rl = (wv * MB + q) * RPM + sub;
      const u32x4 v = chunk(pb[q] + k0, k0 + (slot ^ (rl & 7)) * EPC);
      *reinterpret_cast<RWIG_LDS u32x4 *>(lds + (BM + (wv * MB + q) * RPM) * ROWB + lane * 16) =
          v;
    }
    __syncthreads();
    read(std::integral_constant<int, 0>{}, 0);
    mma(std::integral_constant<int, 0>{});
  };
  if (nst <= 0) {
    if (kend > kbeg) tail(kbeg);
    return;
  }
#pragma unroll
  for (int q = 0; q < STAGES; ++q)
    if (q < nst) issue(q, q);
  wait_later<MOVES, PF>(min(nst, STAGES) - 1);
  barrier();
  read(std::integral_constant<int, 0>{}, 0);
  // Steady step t (stage S = t % STAGES, register set SET = t % 2, both
  // compile-time): no branches, and the scheduling groups below fix the
  // order of the step's MFMAs, moves and fragment reads.
  auto steady = [&](auto s_c, auto set_c, int t) {
    constexpr int S = decltype(s_c)::value, SET = decltype(set_c)::value;
    wait_vmcnt<(PF - 1) * MOVES>();
    wait_lgkmcnt0();
    barrier();
    __builtin_amdgcn_sched_barrier(0);
    issue(S, t + STAGES);
    read(std::integral_constant<int, SET ^ 1>{}, (S + 1) % STAGES);
    mma(set_c);
    constexpr int P1 = MOVES < NM ? MOVES : NM;                // MFMA + move
    constexpr int P2 = NR < NM - P1 ? NR : NM - P1;            // MFMA + read
#pragma unroll
    for (int p = 0; p < P1; ++p) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
      __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load (into LDS)
    }
#pragma unroll
    for (int p = 0; p < P2; ++p) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
    }
    if constexpr (NM - P1 - P2 > 0) __builtin_amdgcn_sched_group_barrier(0x008, NM - P1 - P2, 0);
    if constexpr (NR - P2 > 0) __builtin_amdgcn_sched_group_barrier(0x100, NR - P2, 0);
    if constexpr (MOVES - P1 > 0) __builtin_amdgcn_sched_group_barrier(0x020, MOVES - P1, 0);
    __builtin_amdgcn_sched_barrier(0);
  };
  // the last steps: stages known at run time, waits by the moves left
  auto last = [&](auto set_c, int t) {
    constexpr int SET = decltype(set_c)::value;
    if (t >= nst) return;
    if (t + 1 < nst) {
      wait_later<MOVES, PF - 1>(min(nst, t + STAGES) - (t + 2));
      wait_lgkmcnt0();
      barrier();
      if (t + STAGES < nst) issue(t % STAGES, t + STAGES);
      read(std::integral_constant<int, SET ^ 1>{}, (t + 1) % STAGES);
    }
    mma(set_c);
  };
  // a whole number of stage cycles and register-set pairs per unrolled block
  constexpr int U = STAGES % 2 == 0 ? STAGES : 2 * STAGES;
  int t = 0;
  for (; t + U - 1 + STAGES < nst; t += U)
    unroll_seq(std::make_integer_sequence<int, U>{}, [&](auto i_c) {
      constexpr int I = decltype(i_c)::value;
      steady(std::integral_constant<int, I % STAGES>{}, std::integral_constant<int, I % 2>{},
             t + I);
    });
  for (; t < nst; t += 2) {
    last(std::integral_constant<int, 0>{}, t);
    last(std::integral_constant<int, 1>{}, t + 1);
  }
  if (kbeg + nst * BK < kend) tail(kbeg + nst * BK);
}

template <typename T, class CF, class Op>
__global__ void __launch_bounds__(CF::NT)
k_gemm_tn(int M, int N, int K, const T *__restrict__ A, long lda, const T *__restrict__ B,
          long ldb, const int *__restrict__ nmap, const int *__restrict__ ncount, SplitK<T> sk,
          Op op) {
  if (ncount) {  // column count read on the device (the grid covers an upper bound)
    const int nc = *ncount;
    if ((int)blockIdx.y * CF::BN >= nc) return;
    N = nc;
  }
  using MF = Mfma<T>;
  constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK;
  constexpr int WM = CF::WM, WN = CF::WN;  // per-wave block
  constexpr int TM = CF::TM, TN = CF::TN;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wm = wv % CF::WGM, wn = wv / CF::WGM;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;

  typename MF::acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = MF::zero();

  // k range of this split (multiples of BK)
  const int steps = (K + BK - 1) / BK;
  const int per = (steps + sk.ks - 1) / sk.ks;
  const int kbeg = blockIdx.z * per * BK, kend = min(K, ((int)blockIdx.z + 1) * per * BK);
  if constexpr (CF::DMA)
    gemm_main_dma<T, CF>(M, N, A, lda, B, ldb, nmap, m0, n0, kbeg, kend, acc);
  else
    gemm_main_regs<T, CF>(M, N, K, A, lda, B, ldb, nmap, m0, n0, kbeg, kend, acc);
  const int lr = lane & 15;
  if (sk.work) {
    // partial tile of this split, column-major within the tile; the sum and
    // the epilogue follow in k_gemm_splitk_epilogue (the kernel boundary
    // orders the writes)
    const long ntile = (long)gridDim.x * gridDim.y;
    const long tile = (long)blockIdx.y * gridDim.x + blockIdx.x;
    typename MF::out_t *w = sk.work + ((long)blockIdx.z * ntile + tile) * BM * BN;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int nl = wn * WN + j * 16 + lr, ml = wm * WM + i * 16 + MF::row(lane, r);
          w[nl * BM + ml] = acc[i][j][r];
        }
    return;
  }
  gemm_epilogue<T, CF>(M, N, m0, n0, nmap, acc, op);
}

// Sum of the split partials of each tile, in split order, then the epilogue.
// Workgroup (tile, blockIdx.z) covers 16 columns of the tile with its
// threads along the rows, so that the epilogue's column-major loads and
// stores are contiguous across a wavefront.
constexpr int kEpiCols = 16;

// Index in the split workspace of element (row m, column n) of split z, for a
// product whose grid has gx row tiles and ntile tiles in all: partial tiles
// are column-major within the tile.
template <class CF>
__host__ __device__ inline long tile_partial(int m, int n, int z, int gx, long ntile) {
  const long tile = (long)(n / CF::BN) * gx + m / CF::BM;
  return ((long)z * ntile + tile) * CF::BM * CF::BN + (long)(n % CF::BN) * CF::BM + m % CF::BM;
}

template <typename T, class CF, class Op>
__global__ void __launch_bounds__(256)
k_gemm_splitk_epilogue(int M, int N, const int *__restrict__ nmap,
                       const int *__restrict__ ncount, int ks,
                       const typename Mfma<T>::out_t *__restrict__ work, Op op) {
  constexpr int BM = CF::BM, BN = CF::BN;
  static_assert((BM * kEpiCols) % 256 == 0 && BN % kEpiCols == 0, "epilogue tiling");
  if (ncount) {
    const int nc = *ncount;
    if ((int)blockIdx.y * BN >= nc) return;
    N = nc;
  }
  const long ntile = (long)gridDim.x * gridDim.y;
  const long tile = (long)blockIdx.y * gridDim.x + blockIdx.x;
  const long zs = ntile * BM * BN;
  const typename Mfma<T>::out_t *w0 = work + tile * BM * BN;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
#pragma unroll
  for (int pass = 0; pass < BM * kEpiCols / 256; ++pass) {
    const int idx = pass * 256 + threadIdx.x;
    const int nl = blockIdx.z * kEpiCols + idx / BM, ml = idx % BM;
    const int m = m0 + ml, n = n0 + nl;
    if (m >= M || n >= N) continue;
    const long e = (long)nl * BM + ml;
    typename Mfma<T>::out_t v = w0[e];
    for (int z = 1; z < ks; ++z) v += w0[z * zs + e];
    op(m, nmap ? nmap[n] : n, v);
  }
}

// nmap (optional, N entries): GEMM column n reads column nmap[n] of B and is
// reported to the epilogue as column nmap[n] -- a product over a subset of
// stored columns without moving them. ncount (optional, device): the number
// of columns actually used, read by the kernel (N is then an upper bound that
// sizes the grid), so that a launch can be recorded once in a graph and
// replayed over a changing column set.
// With a split workspace and epilogue = false, only the product runs: the
// caller reads the partial tiles itself (tile_partial gives the index of an
// element), for an update that visits fewer elements than M x N.
template <typename T, class CF = GemmCfg<T>, class Op>
void gemm_tn(hipStream_t st, int M, int N, int K, const T *A, long lda, const T *B, long ldb,
             Op op, const int *nmap = nullptr, const int *ncount = nullptr,
             const SplitK<T> &sk = SplitK<T>(), bool epilogue = true) {
  if constexpr (CF::DMA) {
    // the direct-to-LDS path moves 16-byte chunks of 16-byte aligned rows (a
    // ragged last k-step goes through registers): other operands take the
    // register path with the same tiles (same arithmetic, same split
    // workspace)
    constexpr long EPC = 16 / sizeof(T);
    const bool ok = lda % EPC == 0 && ldb % EPC == 0 &&
                    (reinterpret_cast<uintptr_t>(A) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(B) & 15) == 0;
    if (!ok) {
      using Reg = GemmShape<T, CF::BM, CF::BN, CF::BK, 16, CF::WGM, CF::WGN, 2, 0>;
      gemm_tn<T, Reg>(st, M, N, K, A, lda, B, ldb, op, nmap, ncount, sk, epilogue);
      return;
    }
  }
  constexpr int BM = CF::BM, BN = CF::BN;
  RWIG_REQUIRE(M > 0 && N >= 0 && K > 0, "GEMM dimensions must be positive");
  RWIG_REQUIRE(lda % (16 / sizeof(T)) == 0 && ldb % (16 / sizeof(T)) == 0,
               "GEMM leading dimensions must be multiples of 16 bytes");
  RWIG_REQUIRE(sk.ks >= 1 && (sk.ks == 1 || sk.work), "split-k product needs its workspace");
  if (N == 0) return;
  const dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, sk.ks);
  hipLaunchKernelGGL((k_gemm_tn<T, CF, Op>), grid, dim3(CF::NT), 0, st, M, N, K, A, lda, B,
                     ldb, nmap, ncount, sk, op);
  check_launch();
  if (sk.work && epilogue) {
    hipLaunchKernelGGL((k_gemm_splitk_epilogue<T, CF, Op>), dim3(grid.x, grid.y, BN / kEpiCols),
                       dim3(256), 0, st, M, N, nmap, ncount, sk.ks, sk.work, op);
    check_launch();
  }
}

}  // namespace rwig
