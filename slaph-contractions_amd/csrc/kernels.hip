// Device kernels of the sLapH contraction path (gfx950 / CDNA4).
//
// VdaggerV and the first Q2 stage are dense complex products A^H B with a long
// contraction dimension; they run on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64) with the momentum phase applied while the B tile is
// staged into LDS.  The diluted-factor and trace products are many small
// D x D matrices and run one wavefront per matrix, driven by index tables
// that the host planner builds.
#include "kernels.hpp"

#include <algorithm>
#include <cstdlib>
#include <mutex>

namespace slaph {

// Defined in host.cpp: the sparse gamma tables as plain arrays.
void host_gamma_tables(int (*rows)[4], double (*vals)[4][2]);

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// Factor slabs and product tables are stored as double2 (plain path) or
// float2 (FP32-storage precision form); arithmetic is FP64 either way.
__device__ __forceinline__ double2 ld2(const double2 *p) { return *p; }
__device__ __forceinline__ double2 ld2(const float2 *p) {
  const float2 v = *p;
  return make_double2(double(v.x), double(v.y));
}
__device__ __forceinline__ void st2(double2 *p, double2 v) { *p = v; }
__device__ __forceinline__ void st2(float2 *p, double2 v) { *p = make_float2(float(v.x), float(v.y)); }
// 48-bit form: the dropped 16 bits are rounded half up in magnitude (adding
// 2^15 to the bit pattern carries into the exponent exactly when the kept
// mantissa overflows); finite values only.
__device__ __forceinline__ double2 ld2(const d48c *p) {
  const unsigned a = p->w[0], b = p->w[1], c = p->w[2];
  const unsigned long long re = ((unsigned long long)a << 32) | ((unsigned long long)(c & 0xffffu) << 16);
  const unsigned long long im = ((unsigned long long)b << 32) | ((unsigned long long)(c >> 16) << 16);
  return make_double2(__longlong_as_double((long long)re), __longlong_as_double((long long)im));
}
__device__ __forceinline__ d48c to_d48(double2 v) {
  const unsigned long long re = (unsigned long long)__double_as_longlong(v.x) + 0x8000ull;
  const unsigned long long im = (unsigned long long)__double_as_longlong(v.y) + 0x8000ull;
  d48c o;
  o.w[0] = unsigned(re >> 32);
  o.w[1] = unsigned(im >> 32);
  o.w[2] = unsigned((re >> 16) & 0xffffu) | (unsigned((im >> 16) & 0xffffu) << 16);
  return o;
}
__device__ __forceinline__ void st2(d48c *p, double2 v) { *p = to_d48(v); }
// Streaming (non-temporal) stores for tables that are written once and read
// by a later launch.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st2_nt(double2 *p, double2 v) {
  d2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<d2v *>(p));
}
__device__ __forceinline__ void st2_nt(float2 *p, double2 v) {
  f2v w = {float(v.x), float(v.y)};
  __builtin_nontemporal_store(w, reinterpret_cast<f2v *>(p));
}
// one 12-byte store per element, so that 64 lanes writing consecutive
// elements write 768 contiguous bytes (three 4-byte stores each wrote every
// third word of them)
typedef unsigned u3v __attribute__((ext_vector_type(3)));
__device__ __forceinline__ void st2_nt(d48c *p, double2 v) {
  const d48c o = to_d48(v);
  const u3v w = {o.w[0], o.w[1], o.w[2]};
  __builtin_nontemporal_store(w, reinterpret_cast<u3v *>(p));
}
__device__ __forceinline__ void cfma(double2 &acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// Sparse gamma tables in constant memory (same closed form as the host's
// gamma_table(); uploaded once per process).
__constant__ int c_gamma_row[16][4];
__constant__ double2 c_gamma_val[16][4];

// Lane-permuted copy of a double through DPP (two 32-bit moves); lanes that
// the row mask leaves out read 0.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, kRowMask, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, kRowMask, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes, returned in every lane: pairs, quads (quad_perm),
// 8 and 16 lanes (row half-mirror, row mirror), then rows 0+1 and 2+3
// (row_bcast:15) and all four rows (row_bcast:31) into lane 63.  DPP moves
// are vector-ALU operations, so the chain has no LDS round trips.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_d<0xB1>(v);        // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);        // quad_perm [2,3,0,1]
  v += dpp_d<0x141>(v);       // row_half_mirror
  v += dpp_d<0x140>(v);       // row_mirror
  v += dpp_d<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_d<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------
// A^H diag(phase) B on FP64 MFMA.  Workgroup: 4 waves, 64 x 64 complex output
// tile, each wave a 32 x 32 block = 2 x 2 MFMA 16x16 blocks for the real and
// the imaginary part.  K is staged in chunks of 16 through LDS.
// MFMA 16x16x4 f64 operand layout: A[m=lane%16][k=lane/16],
// B[k=lane/16][n=lane%16]; result element i of lane l is row l/16 + 4*i,
// column l%16 (measured: tests/test_gpu_parity.py::test_ahb_kernel_matches_numpy).
// ---------------------------------------------------------------------------
constexpr int AHB_TM = 64, AHB_TN = 64, AHB_KC = 16;

__global__ __launch_bounds__(256) void k_ahb(const AhBJob *__restrict__ jobs, int M, int N, int K) {
  const AhBJob J = jobs[blockIdx.z];
  const int m0 = blockIdx.x * AHB_TM, n0 = blockIdx.y * AHB_TN;
  __shared__ double2 As[AHB_TM][AHB_KC + 1];
  __shared__ double2 Bs[AHB_TN][AHB_KC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  d4 cr[2][2], ci[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      cr[a][b] = d4{0.0, 0.0, 0.0, 0.0};
      ci[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    }
  for (int k0 = 0; k0 < K; k0 += AHB_KC) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + r * 256;
      const int kk = idx % AHB_KC, mm = idx / AHB_KC;
      const int k = k0 + kk;
      const int m = m0 + mm, n = n0 + mm;
      double2 av = make_double2(0.0, 0.0), bv = make_double2(0.0, 0.0);
      if (k < K && m < M) av = J.A[(long long)m * J.lda + k];
      if (k < K && n < N) {
        bv = J.B[(long long)n * J.ldb + k];
        if (J.phase) bv = cmul(J.phase[k / 3], bv);
      }
      As[mm][kk] = av;
      Bs[mm][kk] = bv;
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < AHB_KC; kq += 4) {
      const int kr = kq + (lane >> 4);
      const double2 a0 = As[wm * 32 + (lane & 15)][kr];
      const double2 a1 = As[wm * 32 + 16 + (lane & 15)][kr];
      const double2 b0 = Bs[wn * 32 + (lane & 15)][kr];
      const double2 b1 = Bs[wn * 32 + 16 + (lane & 15)][kr];
      const double2 av[2] = {a0, a1}, bv[2] = {b0, b1};
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
          // conj(a) b = (ax bx + ay by) + i (ax by - ay bx)
          cr[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mi].x, bv[ni].x, cr[mi][ni], 0, 0, 0);
          cr[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mi].y, bv[ni].y, cr[mi][ni], 0, 0, 0);
          ci[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mi].x, bv[ni].y, ci[mi][ni], 0, 0, 0);
          ci[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[mi].y, bv[ni].x, ci[mi][ni], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = m0 + wm * 32 + mi * 16 + (lane >> 4) + 4 * i;
        const int col = n0 + wn * 32 + ni * 16 + (lane & 15);
        if (row < M && col < N) {
          double2 v = make_double2(cr[mi][ni][i], ci[mi][ni][i]);
          if (J.rowscale) v = cmul(J.rowscale[row], v);
          J.C[(long long)col * J.ldc + row] = v;
        }
      }
}

// Grouped form of the same product for the first Q2 stage: the jobs of rpj
// rows (rpj = D) that share one right matrix B are stacked into one tall
// A^H B, so a 64 x 64 tile holds 64 / D jobs; rows of job j come from
// rows[job0 + j].A and go to rows[job0 + j].C scaled by its rowscale.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_ahb_grouped(const AhBGroup *__restrict__ groups,
                                                     const AhBRow *__restrict__ rows, int rpj, int N, int K,
                                                     long long lda, long long ldb, long long ldc) {
  const AhBGroup G = groups[blockIdx.z];
  const int m0 = blockIdx.x * AHB_TM, n0 = blockIdx.y * AHB_TN;
  const int M = G.njob * rpj;
  if (m0 >= M) return;
  __shared__ double2 As[AHB_TM][AHB_KC + 1];
  __shared__ double2 Bs[AHB_TN][AHB_KC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // this thread's four (row, k) slots of a chunk: the row pointers are
  // resolved once, and the next chunk's elements are loaded into registers
  // while the current one is multiplied
  const double2 *arow[4], *brow[4];
  int kk[4], mm[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int idx = tid + r * 256;
    kk[r] = idx % AHB_KC;
    mm[r] = idx / AHB_KC;
    const int m = m0 + mm[r], n = n0 + mm[r];
    arow[r] = m < M ? rows[G.job0 + m / rpj].A + (long long)(m % rpj) * lda : nullptr;
    brow[r] = n < N ? G.B + (long long)n * ldb : nullptr;
  }
  double2 an[4], bn[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = k0 + kk[r];
      an[r] = k < K && arow[r] ? arow[r][k] : make_double2(0.0, 0.0);
      bn[r] = k < K && brow[r] ? brow[r][k] : make_double2(0.0, 0.0);
    }
  };
  // conj(a) b by three real products (Gauss): P1 = ax bx, P2 = ay by,
  // P3 = (ax - ay)(bx + by); re = P1 + P2, im = P3 - P1 + P2
  d4 p1[2][2], p2[2][2], p3[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) p1[a][b] = p2[a][b] = p3[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  load(0);
  for (int k0 = 0; k0 < K; k0 += AHB_KC) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      As[mm[r]][kk[r]] = an[r];
      Bs[mm[r]][kk[r]] = bn[r];
    }
    __syncthreads();
    if (k0 + AHB_KC < K) load(k0 + AHB_KC);
#pragma unroll
    for (int kq = 0; kq < AHB_KC; kq += 4) {
      const int kr = kq + (lane >> 4);
      const double2 av[2] = {As[wm * 32 + (lane & 15)][kr], As[wm * 32 + 16 + (lane & 15)][kr]};
      const double2 bv[2] = {Bs[wn * 32 + (lane & 15)][kr], Bs[wn * 32 + 16 + (lane & 15)][kr]};
      const double am[2] = {av[0].x - av[0].y, av[1].x - av[1].y};
      const double bp[2] = {bv[0].x + bv[0].y, bv[1].x + bv[1].y};
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
          // operands swapped: the result block is (C column, C row), so that
          // lanes 0..15 of a row group hold 16 consecutive rows of C
          p1[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[ni].x, av[mi].x, p1[mi][ni], 0, 0, 0);
          p2[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[ni].y, av[mi].y, p2[mi][ni], 0, 0, 0);
          p3[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(bp[ni], am[mi], p3[mi][ni], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = m0 + wm * 32 + mi * 16 + (lane & 15);
        const int col = n0 + wn * 32 + ni * 16 + (lane >> 4) + 4 * i;
        if (row < M && col < N) {
          const AhBRow R = rows[G.job0 + row / rpj];
          const double q1 = p1[mi][ni][i], q2 = p2[mi][ni][i];
          const double2 v = cmul(R.rowscale[row % rpj], make_double2(q1 + q2, p3[mi][ni][i] - q1 + q2));
          R.C[(long long)col * ldc + row % rpj] = v;
        }
      }
}

// ---------------------------------------------------------------------------
// VdaggerV for a group of momenta (reference: OperatorFactory.cpp:294-300).
// One workgroup = one 32 x 32 complex output tile of every momentum of the
// group; 4 waves, each a 16 x 16 block.  V is streamed in K-chunks of 32 rows
// through double-buffered LDS (the A and B panels are columns of the same
// V(t), or of V and the displaced W = D V for a displaced operator, whose
// job has one entry and no phase); the momentum phases of the chunk's sites are staged next to them and
// applied to the B fragment in registers, so every V element loaded feeds
// 3 * np MFMAs.  The complex product uses three real products (Gauss):
// conj(a) b = (ax bx + ay by) + i ((ax - ay)(bx + by) - ax bx + ay by), so the
// real and imaginary parts come from P1 = ax bx, P2 = ay by and
// P3 = (ax - ay)(bx + by) accumulated separately: 3 MFMAs instead of 4.
// ---------------------------------------------------------------------------
constexpr int VDV_T = 32, VDV_KC = 32, VDV_SITES = VDV_KC / 3 + 2;

// Operand type, MFMA and output layout per precision: FP64 (the plain path)
// or single-pass FP32 (a precision form: V and the phases rounded to FP32,
// FP32 accumulation).  16x16x4 f64 result element i of lane l sits in row
// l/16 + 4i, the f32 one in row 4(l/16) + i.
template <bool kF32>
struct VdvPrec;
template <>
struct VdvPrec<false> {
  typedef double R;
  typedef double2 S;
  typedef d4 Acc;
  static __device__ Acc mfma(R a, R b, Acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ S cvt(double2 x) { return x; }
  static __device__ int row(int lane, int i) { return (lane >> 4) + 4 * i; }
};
typedef float f4 __attribute__((ext_vector_type(4)));
template <>
struct VdvPrec<true> {
  typedef float R;
  typedef float2 S;
  typedef f4 Acc;
  static __device__ Acc mfma(R a, R b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ S cvt(double2 x) { return make_float2(float(x.x), float(x.y)); }
  static __device__ int row(int lane, int i) { return 4 * (lane >> 4) + i; }
};

template <bool kF32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_vdv(const VdvJob *__restrict__ jobs,
                                                                                      int N, long long K,
                                                                                      int tiles_n) {
  typedef VdvPrec<kF32> P;
  typedef typename P::R R;
  typedef typename P::S S;
  typedef typename P::Acc Acc;
  const VdvJob &J = jobs[blockIdx.y];
  const int np = J.np;
  const int m0 = (blockIdx.x / tiles_n) * VDV_T, n0 = (blockIdx.x % tiles_n) * VDV_T;
  __shared__ S As[2][VDV_T][VDV_KC + 1];
  __shared__ S Bs[2][VDV_T][VDV_KC + 1];
  __shared__ S Ph[2][kVdvGroup][VDV_SITES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const double2 *__restrict__ V = J.V;
  const double2 *__restrict__ Vb = J.W ? J.W : J.V;
  Acc p1[kVdvGroup], p2[kVdvGroup], p3[kVdvGroup];
#pragma unroll
  for (int p = 0; p < kVdvGroup; ++p) {
    p1[p] = Acc{0, 0, 0, 0};
    p2[p] = Acc{0, 0, 0, 0};
    p3[p] = Acc{0, 0, 0, 0};
  }
  double2 ra[4], rb[4], rp = make_double2(1.0, 0.0);
  const int nk = int((K + VDV_KC - 1) / VDV_KC);
  auto load = [&](int kc) {
    const long long k0 = (long long)kc * VDV_KC;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + r * 256;
      const int kk = idx % VDV_KC, c = idx / VDV_KC;
      const long long k = k0 + kk;
      ra[r] = (k < K && m0 + c < N) ? V[(long long)(m0 + c) * K + k] : make_double2(0.0, 0.0);
      rb[r] = (k < K && n0 + c < N) ? Vb[(long long)(n0 + c) * K + k] : make_double2(0.0, 0.0);
    }
    if (tid < kVdvGroup * VDV_SITES) {
      const int p = tid / VDV_SITES, j = tid % VDV_SITES;
      const long long site = k0 / 3 + j;
      rp = make_double2(1.0, 0.0);
      if (p < np && J.phase[p] && site * 3 < K) rp = J.phase[p][site];
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + r * 256;
      As[buf][idx / VDV_KC][idx % VDV_KC] = P::cvt(ra[r]);
      Bs[buf][idx / VDV_KC][idx % VDV_KC] = P::cvt(rb[r]);
    }
    if (tid < kVdvGroup * VDV_SITES) Ph[buf][tid / VDV_SITES][tid % VDV_SITES] = P::cvt(rp);
  };
  load(0);
  store(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) load(kc + 1);
    const long long k0 = (long long)kc * VDV_KC, s0 = k0 / 3;
#pragma unroll
    for (int kq = 0; kq < VDV_KC; kq += 4) {
      const int kr = kq + (lane >> 4);
      const S a = As[buf][wm * 16 + (lane & 15)][kr];
      const S v = Bs[buf][wn * 16 + (lane & 15)][kr];
      const R ad = a.x - a.y;
      const int sl = int((k0 + kr) / 3 - s0);
#pragma unroll
      for (int p = 0; p < kVdvGroup; ++p) {
        if (p < np) {
          const S ph = Ph[buf][p][sl];
          const R bx = ph.x * v.x - ph.y * v.y, by = ph.x * v.y + ph.y * v.x;
          p1[p] = P::mfma(a.x, bx, p1[p]);
          p2[p] = P::mfma(a.y, by, p2[p]);
          p3[p] = P::mfma(ad, bx + by, p3[p]);
        }
      }
    }
    if (kc + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < kVdvGroup; ++p) {
    if (p >= np) continue;
    double2 *C = J.C[p];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = m0 + wm * 16 + P::row(lane, i);
      const int col = n0 + wn * 16 + (lane & 15);
      if (row < N && col < N)
        C[(long long)col * N + row] =
            make_double2(double(p1[p][i]) + double(p2[p][i]), double(p3[p][i]) - double(p1[p][i]) + double(p2[p][i]));
    }
  }
}

// ---------------------------------------------------------------------------
// VdaggerV precision forms on the low-precision matrix cores, one momentum (or
// one displaced operator) per job, same tiling and LDS staging as k_vdv (the
// phase applied to the B panel as it is staged, in FP64).
//  * kForm 0, split BF16: every FP64 operand element x = hi + mid + lo with
//    three BF16 terms (each the RNE rounding of the remainder), products
//    hi.hi + hi.mid + mid.hi + hi.lo + lo.hi + mid.mid on the BF16 MFMA with
//    FP32 accumulation (about 24 significant bits per operand; the K-long
//    FP32 sums bound the accuracy).  Four real products per complex one.
//  * kForm 1, Ozaki int8: each operand column scaled by 2^-e (e from the
//    column's largest modulus, so that |x| < 1) and cut into kOzakiSlices
//    int8 slices of 7 bits, x = sum_i s_i 2^-7i; every slice pair with
//    i + j <= S + 1 multiplied exactly on the int8 MFMA into int32 sums per
//    diagonal d = i + j, which are moved into FP64 accumulators (times
//    2^-7d) every kOzakiFlush K-chunks, before an int32 sum can overflow
//    (2 S 127^2 32 kOzakiFlush < 2^31); the result is scaled back by
//    2^(eA + eB).  Only the dropped slice pairs and the truncation of the
//    slicing limit the accuracy: about 2^-7(S+1) of max|A| max|B| K.
// Lane layouts (checked on the device): bf16 16x16x16 A lane l row l%16,
// k = 4(l/16) + j; i8 16x16x32 k = 8(l/16) + j; results row 4(l/16) + i,
// column l%16.
// ---------------------------------------------------------------------------
typedef int i4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));
constexpr int kOzakiFlush = 256;

__device__ __forceinline__ short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return short(u >> 16);
}
__device__ __forceinline__ double bf16_value(short h) { return double(__uint_as_float(unsigned(static_cast<unsigned short>(h)) << 16)); }

// three-term BF16 split of x: x = hi + mid + lo up to the last term's rounding
__device__ __forceinline__ void split3(double x, short &hi, short &mid, short &lo) {
  hi = bf16_rne(float(x));
  const double r1 = x - bf16_value(hi);
  mid = bf16_rne(float(r1));
  const double r2 = r1 - bf16_value(mid);
  lo = bf16_rne(float(r2));
}

template <int S>
__device__ __forceinline__ void ozaki_slices(const double (&x)[8], long long (&out)[S]) {
#pragma unroll
  for (int i = 0; i < S; ++i) out[i] = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    double v = x[j];
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const double y = v * 128.0;
      const double t = trunc(y);  // |t| <= 127 since |v| < 1
      v = y - t;                 // exact
      out[i] |= (long long)(unsigned char)(signed char)int(t) << (8 * j);
    }
  }
}

template <int kForm>
__global__ __launch_bounds__(256) void k_vdv_split(const VdvJob *__restrict__ jobs, int N, long long K, int tiles_n) {
  constexpr int S = kOzakiSlices;
  const VdvJob &J = jobs[blockIdx.y];
  const int m0 = (blockIdx.x / tiles_n) * VDV_T, n0 = (blockIdx.x % tiles_n) * VDV_T;
  __shared__ double2 As[2][VDV_T][VDV_KC + 1];
  __shared__ double2 Bs[2][VDV_T][VDV_KC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const double2 *__restrict__ V = J.V;
  const double2 *__restrict__ Vb = J.W ? J.W : J.V;
  const double2 *__restrict__ ph = J.phase[0];
  double2 ra[4], rb[4];
  const int nk = int((K + VDV_KC - 1) / VDV_KC);
  auto load = [&](int kc) {
    const long long k0 = (long long)kc * VDV_KC;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + r * 256;
      const int kk = idx % VDV_KC, c = idx / VDV_KC;
      const long long k = k0 + kk;
      ra[r] = (k < K && m0 + c < N) ? V[(long long)(m0 + c) * K + k] : make_double2(0.0, 0.0);
      double2 b = (k < K && n0 + c < N) ? Vb[(long long)(n0 + c) * K + k] : make_double2(0.0, 0.0);
      if (ph && k < K) {
        const double2 p = ph[k / 3];
        b = make_double2(p.x * b.x - p.y * b.y, p.x * b.y + p.y * b.x);
      }
      rb[r] = b;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + r * 256;
      As[buf][idx / VDV_KC][idx % VDV_KC] = ra[r];
      Bs[buf][idx / VDV_KC][idx % VDV_KC] = rb[r];
    }
  };
  const int arow = wm * 16 + (lane & 15), brow = wn * 16 + (lane & 15);
  // per-column scales of the Ozaki form (2^-e of this lane's A row / B column)
  double sa = 1.0, sb = 1.0;
  int ea = 0, eb = 0;
  if (kForm == 1) {
    if (m0 + arow < N) ea = J.ex[0][m0 + arow];
    if (n0 + brow < N) eb = J.ex[1][n0 + brow];
    sa = ldexp(1.0, -ea);
    sb = ldexp(1.0, -eb);
  }
  f4 facc_re = {0, 0, 0, 0}, facc_im = {0, 0, 0, 0};  // split BF16
  i4v iacc_re[S], iacc_im[S];                          // Ozaki, diagonal d = 2 .. S + 1
  double dre[4] = {0, 0, 0, 0}, dim[4] = {0, 0, 0, 0};
#pragma unroll
  for (int d = 0; d < S; ++d) {
    iacc_re[d] = i4v{0, 0, 0, 0};
    iacc_im[d] = i4v{0, 0, 0, 0};
  }
  auto flush = [&]() {
#pragma unroll
    for (int d = 0; d < S; ++d) {
      const double sc = ldexp(1.0, -7 * (d + 2));
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        dre[i] += sc * double(iacc_re[d][i]);
        dim[i] += sc * double(iacc_im[d][i]);
      }
      iacc_re[d] = i4v{0, 0, 0, 0};
      iacc_im[d] = i4v{0, 0, 0, 0};
    }
  };
  load(0);
  store(0);
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) load(kc + 1);
    if (kForm == 0) {
#pragma unroll
      for (int kq = 0; kq < VDV_KC; kq += 16) {
        s4v axh, axm, axl, ayh, aym, ayl, bxh, bxm, bxl, byh, bym, byl;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = kq + 4 * (lane >> 4) + j;
          const double2 a = As[buf][arow][k], b = Bs[buf][brow][k];
          short h, m, l;
          split3(a.x, h, m, l);
          axh[j] = h, axm[j] = m, axl[j] = l;
          split3(a.y, h, m, l);
          ayh[j] = h, aym[j] = m, ayl[j] = l;
          split3(b.x, h, m, l);
          bxh[j] = h, bxm[j] = m, bxl[j] = l;
          split3(b.y, h, m, l);
          byh[j] = h, bym[j] = m, byl[j] = l;
        }
        // conj(a) b: Re = ax bx + ay by, Im = ax by - ay bx
        auto prod = [&](f4 c, s4v xh, s4v xm, s4v xl, s4v yh, s4v ym, s4v yl) {
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xl, yh, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xh, yl, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xm, ym, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xm, yh, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xh, ym, c, 0, 0, 0);
          return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xh, yh, c, 0, 0, 0);
        };
        facc_re = prod(facc_re, axh, axm, axl, bxh, bxm, bxl);
        facc_re = prod(facc_re, ayh, aym, ayl, byh, bym, byl);
        facc_im = prod(facc_im, axh, axm, axl, byh, bym, byl);
        const s4v sg = {short(0x8000), short(0x8000), short(0x8000), short(0x8000)};
        facc_im = prod(facc_im, ayh ^ sg, aym ^ sg, ayl ^ sg, bxh, bxm, bxl);  // -ay: sign bits flipped
      }
    } else {
      double ax[8], ay[8], bx[8], by[8], nbx[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = 8 * (lane >> 4) + j;
        const double2 a = As[buf][arow][k], b = Bs[buf][brow][k];
        ax[j] = a.x * sa;
        ay[j] = a.y * sa;
        bx[j] = b.x * sb;
        by[j] = b.y * sb;
        nbx[j] = -bx[j];
      }
      long long sx[S], sy[S], tx[S], ty[S], tn[S];
      ozaki_slices<S>(ax, sx);
      ozaki_slices<S>(ay, sy);
      ozaki_slices<S>(bx, tx);
      ozaki_slices<S>(by, ty);
      ozaki_slices<S>(nbx, tn);
#pragma unroll
      for (int d = 0; d < S; ++d)      // diagonal i + j = d + 2 (slices numbered from 1)
#pragma unroll
        for (int i = 0; i <= d; ++i) {  // slice i + 1 of A with slice d - i + 1 of B
          const int j = d - i;
          iacc_re[d] = __builtin_amdgcn_mfma_i32_16x16x32_i8(sx[i], tx[j], iacc_re[d], 0, 0, 0);
          iacc_re[d] = __builtin_amdgcn_mfma_i32_16x16x32_i8(sy[i], ty[j], iacc_re[d], 0, 0, 0);
          iacc_im[d] = __builtin_amdgcn_mfma_i32_16x16x32_i8(sx[i], ty[j], iacc_im[d], 0, 0, 0);
          iacc_im[d] = __builtin_amdgcn_mfma_i32_16x16x32_i8(sy[i], tn[j], iacc_im[d], 0, 0, 0);
        }
      if ((kc + 1) % kOzakiFlush == 0) flush();
    }
    if (kc + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
  if (kForm == 1) flush();
  double2 *C = J.C[0];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = m0 + wm * 16 + 4 * (lane >> 4) + i;
    const int col = n0 + wn * 16 + (lane & 15);
    if (row < N && col < N) {
      double2 v;
      if (kForm == 0) {
        v = make_double2(double(facc_re[i]), double(facc_im[i]));
      } else {
        // this lane's A scale belongs to row wm*16 + lane%16, not to `row`
        const int er = J.ex[0][row], ec = J.ex[1][col];
        v = make_double2(ldexp(dre[i], er + ec), ldexp(dim[i], er + ec));
      }
      C[(long long)col * N + row] = v;
    }
  }
}

// C = A^H B for tall K (displaced VdaggerV: A = V(t), B = D V(t), K = 3 L^3):
// 64 x 64 output tile per workgroup, 4 waves of 32 x 32 (2 x 2 MFMA tiles,
// three real products per complex one: Re = p1 + p2, Im = p3 - p1 + p2 for
// conj(a) b), K split into `parts` ranges of whole 16-row chunks staged
// through double-buffered LDS; each part writes its partial tile, and
// k_ahb_big_reduce sums the parts in order (no atomics: the result does not
// depend on the schedule).
constexpr int kBigT = 64, kBigKC = 16;

// Two waves per SIMD (195 VGPRs, no spills; 2 x 68 KiB of LDS per CU): 22 ms
// per 48^3 x 96 timeslice of three N = 400 products against 31 ms at one.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_ahb_big(
    const AhbBigJob *__restrict__ jobs, int N, long long K, int tiles, long long kper, double2 *__restrict__ part) {
  const AhbBigJob J = jobs[blockIdx.z];
  const int tile = blockIdx.x, p = blockIdx.y, parts = gridDim.y;
  const int m0 = (tile / tiles) * kBigT, n0 = (tile % tiles) * kBigT;
  const long long kb = (long long)p * kper, ke = min(K, kb + kper);
  __shared__ double2 As[2][kBigT][kBigKC + 1];
  __shared__ double2 Bs[2][kBigT][kBigKC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  d4 acc[2][2][3];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[a][b][k] = d4{0.0, 0.0, 0.0, 0.0};
  double2 ra[4], rb[4];
  const int nk = int((ke - kb + kBigKC - 1) / kBigKC);
  auto load = [&](int kc) {
    const long long k0 = kb + (long long)kc * kBigKC;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + r * 256, kk = idx % kBigKC, c = idx / kBigKC;
      const long long k = k0 + kk;
      ra[r] = (k < ke && m0 + c < N) ? J.A[(long long)(m0 + c) * K + k] : make_double2(0.0, 0.0);
      rb[r] = (k < ke && n0 + c < N) ? J.B[(long long)(n0 + c) * K + k] : make_double2(0.0, 0.0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + r * 256;
      As[buf][idx / kBigKC][idx % kBigKC] = ra[r];
      Bs[buf][idx / kBigKC][idx % kBigKC] = rb[r];
    }
  };
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kc = 0; kc < nk; ++kc) {
    const int buf = kc & 1;
    if (kc + 1 < nk) load(kc + 1);
#pragma unroll
    for (int kq = 0; kq < kBigKC; kq += 4) {
      const int kr = kq + (lane >> 4), l16 = lane & 15;
      const double2 a[2] = {As[buf][wm * 32 + l16][kr], As[buf][wm * 32 + 16 + l16][kr]};
      const double2 b[2] = {Bs[buf][wn * 32 + l16][kr], Bs[buf][wn * 32 + 16 + l16][kr]};
#pragma unroll
      for (int ta = 0; ta < 2; ++ta) {
        const double ad = a[ta].x - a[ta].y;
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
          acc[ta][tb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta].x, b[tb].x, acc[ta][tb][0], 0, 0, 0);
          acc[ta][tb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta].y, b[tb].y, acc[ta][tb][1], 0, 0, 0);
          acc[ta][tb][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, b[tb].x + b[tb].y, acc[ta][tb][2], 0, 0, 0);
        }
      }
    }
    if (kc + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
  // partial tile, column-major 64 x 64: [job][part][tile]
  double2 *o = part + (((long long)blockIdx.z * parts + p) * gridDim.x + tile) * (kBigT * kBigT);
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = wm * 32 + ta * 16 + (lane >> 4) + 4 * i, col = wn * 32 + tb * 16 + (lane & 15);
        const d4 &q1 = acc[ta][tb][0], &q2 = acc[ta][tb][1], &q3 = acc[ta][tb][2];
        o[col * kBigT + row] = make_double2(q1[i] + q2[i], q3[i] - q1[i] + q2[i]);
      }
}

__global__ void k_ahb_big_reduce(const AhbBigJob *__restrict__ jobs, int njobs, int N, int tiles, int parts,
                                 const double2 *__restrict__ part) {
  const long long total = (long long)njobs * N * N;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const int job = int(id / ((long long)N * N));
    const int e = int(id % ((long long)N * N)), row = e % N, col = e / N;
    const int tile = (row / kBigT) * tiles + col / kBigT;
    const int local = (col % kBigT) * kBigT + row % kBigT;
    double2 s = make_double2(0.0, 0.0);
    for (int p = 0; p < parts; ++p) {
      const double2 v = part[(((long long)job * parts + p) * tiles * tiles + tile) * (kBigT * kBigT) + local];
      s.x += v.x;
      s.y += v.y;
    }
    jobs[job].C[(long long)col * N + row] = s;
  }
}

// e[n] with max_k |V[k, n]| (1 + 2^-40) < 2^e per column of a K x N matrix
// (one workgroup per column); 0 for a zero column.
__global__ __launch_bounds__(256) void k_colexp(const double2 *__restrict__ V, long long K, int *__restrict__ e) {
  const int n = blockIdx.x;
  double m = 0.0;
  for (long long k = threadIdx.x; k < K; k += 256) {
    const double2 v = V[(long long)n * K + k];
    m = fmax(m, hypot(v.x, v.y));
  }
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int ex = 0;
    if (red[0] > 0.0) frexp(red[0] * (1.0 + 0x1p-40), &ex);
    e[n] = ex;
  }
}

// ---------------------------------------------------------------------------
// VdaggerV by the sitewise Fourier factorisation (see kernels.hpp).  One
// single-wave workgroup = one (timeslice, x, tile pair I <= J).  Per site the
// wave forms W = V_I^H V_J on the 16 x 16 tile by four accumulating MFMAs
// (lane l: colour l/16, column l%16; colour 3 is zero padding), Re W =
// ax bx + ay by and Im W = ax by - ay bx, and transforms all four accumulator
// rows, so that the z sums (cos / sin per |pz|) and the y sums per (py, pz)
// stay in registers.  Per site: 4 MFMAs and 4 (2 + 4 nk) + 3 FP64 vector
// operations per lane.  The operands of the next four sites are loaded while
// the current four are processed.
// ---------------------------------------------------------------------------
template <int NC, int NK, bool kFast>
__device__ __forceinline__ void vdv_ft_wave(const double2 *__restrict__ Vr, FtLayout lay, int I, int J, int Ly, int Lz,
                                            long long site0, FtPass const &P, const double2 *__restrict__ ey,
                                            const double2 *zl, double2 *__restrict__ Uo, long long ustride) {
  const int lane = threadIdx.x & 63, c = lane >> 4, m = lane & 15;
  // Colour-3 lanes (the fourth K slot) read colour 2 and zero their A
  // operand, so that every load is unconditional.
  const double amask = c < 3 ? 1.0 : 0.0;
  const int cl = c < 3 ? c : 2;
  const long long sstep = 3LL * lay.kstride;
  const double2 *pa = Vr + (site0 * 3 + cl) * lay.kstride + (long long)min(I * 16 + m, lay.ncol - 1) * lay.nstride;
  const double2 *pb = Vr + (site0 * 3 + cl) * lay.kstride + (long long)min(J * 16 + m, lay.ncol - 1) * lay.nstride;
  const double2 zero2 = make_double2(0.0, 0.0);
  double2 U[NC][4];
#pragma unroll
  for (int k = 0; k < NC; ++k)
#pragma unroll
    for (int q = 0; q < 4; ++q) U[k][q] = zero2;
  double a0r[4], a0i[4], cr[NK][4], ci[NK][4], sr[NK][4], si[NK][4];
  // W of one site: wr = Re W, wi = Im W.  The four-MFMA form leaves no
  // vector work between the MFMAs and the z transform (the three-multiply
  // form needs three additions per element, and the wave is bound by its
  // vector instructions).
  auto mf = [&](const double2 &a, const double2 &b, d4 &wr, d4 &wi) {
    const d4 dz = {0.0, 0.0, 0.0, 0.0};
    const double ax = a.x * amask, ay = a.y * amask, nay = a.y * -amask;
    wr = __builtin_amdgcn_mfma_f64_16x16x4f64(ax, b.x, dz, 0, 0, 0);
    wi = __builtin_amdgcn_mfma_f64_16x16x4f64(ax, b.y, dz, 0, 0, 0);
    wr = __builtin_amdgcn_mfma_f64_16x16x4f64(ay, b.y, wr, 0, 0, 0);
    wi = __builtin_amdgcn_mfma_f64_16x16x4f64(nay, b.x, wi, 0, 0, 0);
  };
  // z transform of the four rows, site phases cs (cos, sin).
  auto zacc = [&](const double2 *cs, const d4 &w_r, const d4 &w_i) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double wr = w_r[q], wi = w_i[q];
      a0r[q] += wr;
      a0i[q] += wi;
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        cr[k][q] = fma(cs[k].x, wr, cr[k][q]);
        ci[k][q] = fma(cs[k].x, wi, ci[k][q]);
        sr[k][q] = fma(cs[k].y, wr, sr[k][q]);
        si[k][q] = fma(cs[k].y, wi, si[k][q]);
      }
    }
  };
  auto zero_z = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a0r[q] = a0i[q] = 0.0;
#pragma unroll
      for (int k = 0; k < NK; ++k) cr[k][q] = ci[k][q] = sr[k][q] = si[k][q] = 0.0;
    }
  };
  // End of a z line: T(+-k) = (CR +- SI) + i (CI -+ SR), then
  // U[combination] += exp(-i 2 pi py y / Ly) T(pz).
  auto yline = [&](int y) {
#pragma unroll
    for (int k2 = 0; k2 < NC; ++k2) {
      if (k2 < P.nc) {
        const int sel = P.tsel[k2];
        const double2 e = ey[k2 * Ly + y];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double tr = a0r[q], ti = a0i[q];
#pragma unroll
          for (int k = 0; k < NK; ++k) {
            if (sel == 1 + 2 * k) {
              tr = cr[k][q] + si[k][q];
              ti = ci[k][q] - sr[k][q];
            } else if (sel == 2 + 2 * k) {
              tr = cr[k][q] - si[k][q];
              ti = ci[k][q] + sr[k][q];
            }
          }
          U[k2][q].x = fma(e.x, tr, fma(-e.y, ti, U[k2][q].x));
          U[k2][q].y = fma(e.x, ti, fma(e.y, tr, U[k2][q].y));
        }
      }
    }
  };
  double2 na[4], nb[4];
  if constexpr (kFast) {
    double2 nc_[4][NK];
    // Lz % 4 == 0: a group is four consecutive sites of one z line; the
    // operand pointers and the z position advance incrementally.
    const int ng = Ly * (Lz >> 2);
    const long long gstep = 4 * sstep;
    const double2 *qa = pa, *qb = pb;
    auto load = [&](int z0) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        na[u] = qa[u * sstep];
        nb[u] = qb[u * sstep];
#pragma unroll
        for (int k = 0; k < NK; ++k) nc_[u][k] = zl[k * Lz + z0 + u];
      }
    };
    load(0);
    int y = 0, z0 = 0;
    for (int g = 0; g < ng; ++g) {
      double2 ca[4], cb[4], cs[4][NK];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ca[u] = na[u];
        cb[u] = nb[u];
#pragma unroll
        for (int k = 0; k < NK; ++k) cs[u][k] = nc_[u][k];
      }
      const int zn = z0 + 4 == Lz ? 0 : z0 + 4;
      if (g + 1 < ng) {
        qa += gstep;
        qb += gstep;
        load(zn);
      }
      if (z0 == 0) zero_z();
      // Software pipeline: the MFMAs of site u + 1 are issued before the
      // vector work of site u.
      d4 e1, e2, o1, o2;
      mf(ca[0], cb[0], e1, e2);
      mf(ca[1], cb[1], o1, o2);
      zacc(cs[0], e1, e2);
      mf(ca[2], cb[2], e1, e2);
      zacc(cs[1], o1, o2);
      mf(ca[3], cb[3], o1, o2);
      zacc(cs[2], e1, e2);
      zacc(cs[3], o1, o2);
      if (zn == 0) {
        yline(y);
        ++y;
      }
      z0 = zn;
    }
  } else {
    const int GZ = (Lz + 3) >> 2, ng = Ly * GZ;
    auto load = [&](int g) {
      const int y = g / GZ, z0 = (g - y * GZ) * 4;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const long long off = (long long)(y * Lz + min(z0 + u, Lz - 1)) * sstep;
        na[u] = pa[off];
        nb[u] = pb[off];
      }
    };
    load(0);
    for (int g = 0; g < ng; ++g) {
      double2 ca[4], cb[4], cs[4][NK];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ca[u] = na[u];
        cb[u] = nb[u];
      }
      if (g + 1 < ng) load(g + 1);
      const int y = g / GZ, z0 = (g - y * GZ) * 4;
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < NK; ++k) cs[u][k] = zl[k * Lz + min(z0 + u, Lz - 1)];
      if (z0 == 0) zero_z();
      d4 e1, e2, o1, o2;
      mf(ca[0], cb[0], e1, e2);
      if (z0 + 1 < Lz) mf(ca[1], cb[1], o1, o2);
      zacc(cs[0], e1, e2);
      if (z0 + 2 < Lz) mf(ca[2], cb[2], e1, e2);
      if (z0 + 1 < Lz) zacc(cs[1], o1, o2);
      if (z0 + 3 < Lz) mf(ca[3], cb[3], o1, o2);
      if (z0 + 2 < Lz) zacc(cs[2], e1, e2);
      if (z0 + 3 < Lz) zacc(cs[3], o1, o2);
      if (z0 + 4 >= Lz) yline(y);
    }
  }
  // Accumulator row r of lane l is tile row l/16 + 4r, column l%16: pair
  // index (l/16 + 4r) * 16 + l%16 = l + 64 r.
#pragma unroll
  for (int k2 = 0; k2 < NC; ++k2) {
    if (k2 < P.nc) {
#pragma unroll
      for (int q = 0; q < 4; ++q) Uo[k2 * ustride + lane + 64 * q] = U[k2][q];
    }
  }
}

// The accumulators of four rows need more than 256 registers per lane: one
// wave per SIMD.
template <int NC, int NK, bool kFast>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_vdv_ft(
    const double2 *__restrict__ Vr, long long vstride, FtLayout lay, int Lx, int Ly, int Lz,
    const int2 *__restrict__ tiles, int ntp, FtPass P, const double2 *__restrict__ ey,
    const double2 *__restrict__ zcs, double2 *__restrict__ U) {
  // XCD x runs the x-th contiguous eighth of the (t, x, tile pair) blocks, so
  // the tile pairs of one x slab share that XCD's L2.
  const int G = gridDim.x, per = G / 8, rem = G % 8, xcd = blockIdx.x % 8, jx = blockIdx.x / 8;
  const int lb = xcd < rem ? xcd * (per + 1) + jx : rem * (per + 1) + (xcd - rem) * per + jx;
  const int tp = lb % ntp, x = (lb / ntp) % Lx, t = lb / (ntp * Lx);
  const int2 ij = tiles[tp];
  const long long site0 = (long long)x * Ly * Lz;
  const long long ustride = (long long)ntp * 256;
  double2 *Uo = U + (((long long)t * Lx + x) * P.nc * ntp + tp) * 256;
  // z phases of the pass in LDS (read with one broadcast per site and |pz|)
  __shared__ double2 zl[NK * kFtMaxLz];
  for (int i = threadIdx.x; i < NK * Lz; i += blockDim.x) zl[(i / Lz) * Lz + i % Lz] = zcs[(i / Lz) * Lz + i % Lz];
  __syncthreads();
  const double2 *V = Vr + t * vstride;
  vdv_ft_wave<NC, NK, kFast>(V, lay, ij.x, ij.y, Ly, Lz, site0, P, ey, zl, Uo, ustride);
}

// C(p) from the per-x sums: thread = (row, column) of one tile pair.  The
// targets come sorted by combination: each run of up to kXsumGroup targets
// that read the same combination's sums loads every per-x value once and
// feeds all of them (each target's sum over x in ascending x, as alone).
constexpr int kXsumGroup = 8;
__global__ __launch_bounds__(256) void k_vdv_xsum(const double2 *__restrict__ U, int nc, int ntp, int Lx,
                                                  const int2 *__restrict__ tiles, const FtTarget *__restrict__ tg,
                                                  int ntg, const double2 *__restrict__ ex, int N, int t0,
                                                  long long tstride) {
  const int tp = blockIdx.x, t = blockIdx.y, q = threadIdx.x;
  const int row = q & 15, col = q >> 4;
  const int2 ij = tiles[tp];
  const int i = ij.x * 16 + row, j = ij.y * 16 + col;
  const bool ok = i < N && j < N;
  const long long xstride = (long long)nc * ntp * 256;
  const double2 *u0 = U + (long long)t * Lx * xstride + (long long)tp * 256 + row * 16 + col;
  for (int k0 = 0; k0 < ntg;) {
    const int c = tg[k0].c;
    int k1 = k0 + 1;
    while (k1 < ntg && k1 - k0 < kXsumGroup && tg[k1].c == c) ++k1;
    const int ng = k1 - k0;
    const double2 *u = u0 + (long long)c * ntp * 256;
    double sr[kXsumGroup], si[kXsumGroup];
#pragma unroll
    for (int g = 0; g < kXsumGroup; ++g) sr[g] = si[g] = 0.0;
    for (int x = 0; x < Lx; ++x) {
      const double2 v = u[x * xstride];
#pragma unroll
      for (int g = 0; g < kXsumGroup; ++g) {
        if (g < ng) {
          const double2 f = ex[(long long)(k0 + g) * Lx + x];
          sr[g] = fma(f.x, v.x, fma(-f.y, v.y, sr[g]));
          si[g] = fma(f.x, v.y, fma(f.y, v.x, si[g]));
        }
      }
    }
#pragma unroll
    for (int g = 0; g < kXsumGroup; ++g) {
      if (g >= ng || !ok) continue;
      const FtTarget T = tg[k0 + g];
      if (T.lower && ij.x == ij.y) continue;
      double2 *C = T.C + (long long)(t0 + t) * tstride;
      if (!T.lower)
        C[(long long)j * N + i] = make_double2(sr[g], si[g]);
      else
        C[(long long)i * N + j] = make_double2(sr[g], -si[g]);
    }
    k0 = k1;
  }
}

// Column-major K x N -> row-major K x Npad through a 64 x 16 LDS tile.
__global__ __launch_bounds__(256) void k_vrow(const double2 *__restrict__ V, double2 *__restrict__ Vr, long long K,
                                              int N, int Npad) {
  __shared__ double2 tile[16][65];
  const long long k0 = (long long)blockIdx.x * 64;
  const int n0 = blockIdx.y * 16, t = blockIdx.z;
  const double2 *src = V + (long long)t * K * N;
  double2 *dst = Vr + (long long)t * K * Npad;
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int kk = tid & 63, cc = (tid >> 6) + 4 * r;
    const long long k = k0 + kk;
    tile[cc][kk] = (k < K && n0 + cc < N) ? src[(long long)(n0 + cc) * K + k] : make_double2(0.0, 0.0);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int cc = tid & 15, kk = (tid >> 4) + 16 * r;
    const long long k = k0 + kk;
    if (k < K) dst[k * Npad + n0 + cc] = tile[cc][kk];
  }
}

// ---------------------------------------------------------------------------
// Q0[(e1 + dilE g(d)), (e2 + dilE d)] =
//   gval(d) sum_{v=e1 mod dilE} conj rho1(v, g(d)) sum_{w=e2 mod dilE} VdV[v,w] rho2(w,d)
// (reference: DilutedFactorFactory<Q0>::build, DilutedFactorFactory.cpp:33-101)
// ---------------------------------------------------------------------------
// One workgroup per group of jobs that share VdV, rho2 and gamma: X is formed
// once in LDS, then every job of the group reads it with its own rho1.
__global__ __launch_bounds__(256) void k_q0(const Q0Group *__restrict__ groups, const Q0Out *__restrict__ outs,
                                            int N, int dilE) {
  extern __shared__ double2 X[];  // [v][d][e2] : N * 4 * dilE
  const Q0Group G = groups[blockIdx.x];
  const int D = 4 * dilE;
  for (int idx = threadIdx.x; idx < N * dilE; idx += blockDim.x) {
    const int v = idx % N, e2 = idx / N;
    double2 acc[4];
    for (int d = 0; d < 4; ++d) acc[d] = make_double2(0.0, 0.0);
    for (int w = e2; w < N; w += dilE) {
      const double2 x = G.vdv[(long long)w * N + v];
#pragma unroll
      for (int d = 0; d < 4; ++d) cfma(acc[d], x, G.rho2[4 * w + d]);
    }
    for (int d = 0; d < 4; ++d) X[(v * 4 + d) * dilE + e2] = acc[d];
  }
  __syncthreads();
  for (int jj = 0; jj < G.njob; ++jj) {
    const Q0Out O = outs[G.job_off + jj];
    for (int o = threadIdx.x; o < D * D; o += blockDim.x) {
      const int row = o % D, col = o / D;
      const int d = col / dilE, e2 = col % dilE;
      const int g = c_gamma_row[G.gamma][d];
      double2 s = make_double2(0.0, 0.0);
      if (row / dilE == g) {
        const int e1 = row % dilE;
        for (int v = e1; v < N; v += dilE)
          cfma(s, make_double2(O.rho1[4 * v + g].x, -O.rho1[4 * v + g].y), X[(v * 4 + d) * dilE + e2]);
        s = cmul(c_gamma_val[G.gamma][d], s);
      }
      O.out[o] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Q1[(r,e1),(c,e2)] = gval(r) sum_{v=e1} conj rho1(v,r) sum_w VdV[v,w]
//                      P2[(t,g(r),w),(b,c,e2)]
// (reference: DilutedFactorFactory<Q1>::build, DilutedFactorFactory.cpp:103-170)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_q1(const Q1Job *__restrict__ jobs, int N, int dilE, long long prow) {
  extern __shared__ double2 RV[];  // [w][d][e1]
  const Q1Job J = jobs[blockIdx.x];
  const int D = 4 * dilE;
  for (int idx = threadIdx.x; idx < N * dilE; idx += blockDim.x) {
    const int e1 = idx % dilE, w = idx / dilE;
    double2 acc[4];
    for (int d = 0; d < 4; ++d) acc[d] = make_double2(0.0, 0.0);
    for (int v = e1; v < N; v += dilE) {
      const double2 x = J.vdv[(long long)w * N + v];
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const double2 r = J.rho1[4 * v + d];
        cfma(acc[d], make_double2(r.x, -r.y), x);
      }
    }
    for (int d = 0; d < 4; ++d) RV[(w * 4 + d) * dilE + e1] = acc[d];
  }
  __syncthreads();
  for (int o = threadIdx.x; o < D * D; o += blockDim.x) {
    const int row = o % D, col = o / D;
    const int r = row / dilE, e1 = row % dilE;
    const int g = c_gamma_row[J.gamma][r];
    const double2 *pcol = J.peram2 + (long long)(J.b * D + col) * prow + (long long)(J.t * 4 + g) * N;
    double2 s = make_double2(0.0, 0.0);
    for (int w = 0; w < N; ++w) cfma(s, RV[(w * 4 + r) * dilE + e1], pcol[w]);
    J.out[o] = cmul(c_gamma_val[J.gamma][r], s);
  }
}

// ---------------------------------------------------------------------------
// Q2 second stage: Q2[(c,e),(c2,e2)] = sum_bd sum_w M'[(c,e),(bd,w)]
//   P2[(t, g(bd), w), (b2, c2, e2)], M' already carries gval(bd) and the
//   gamma_5 signs (reference: DilutedFactorFactory.cpp:221-243).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_q2b(const Q2BJob *__restrict__ jobs, int N, int dilE, long long prow) {
  const Q2BJob J = jobs[blockIdx.x];
  const int D = 4 * dilE;
  for (int o = threadIdx.x; o < D * D; o += blockDim.x) {
    const int row = o % D, col = o / D;
    double2 s = make_double2(0.0, 0.0);
    for (int bd = 0; bd < 4; ++bd) {
      const int g = c_gamma_row[J.gamma][bd];
      const double2 *pcol = J.peram2 + (long long)(J.b2 * D + col) * prow + (long long)(J.t * 4 + g) * N;
      const double2 *mrow = J.Mp + (long long)bd * N * D + row;
      for (int w = 0; w < N; ++w) cfma(s, mrow[(long long)w * D], pcol[w]);
    }
    J.out[o] = s;
  }
}

// D = 16: Q2 = M' (16 x 4N) * stacked P2 rows (4N x 16) as one MFMA tile per
// job (one wavefront): K index k = bd*N + w; lane l supplies M'[l%16][4s + l/16]
// (address k*16 + row: 1 KB coalesced per k-step) and P2[4s + l/16][l%16]
// (4 consecutive w of one perambulator column).  4 jobs per workgroup.
__global__ __launch_bounds__(256) void k_q2b16(const Q2BJob *__restrict__ jobs, int njobs, int N, long long prow) {
  const int job = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (job >= njobs) return;
  const Q2BJob J = jobs[job];
  const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15;
  // (ax + i ay)(bx + i by) by three real products (Gauss): P1 = ax bx,
  // P2 = ay by, P3 = (ax + ay)(bx + by); re = P1 - P2, im = P3 - P1 - P2
  d4 p1 = d4{0.0, 0.0, 0.0, 0.0}, p2 = p1, p3 = p1;
  for (int bd = 0; bd < 4; ++bd) {
    const int g = c_gamma_row[J.gamma][bd];
    const double2 *pcol = J.peram2 + (long long)(J.b2 * 16 + col) * prow + (long long)(J.t * 4 + g) * N + q;
    const double2 *mrow = J.Mp + (long long)bd * N * 16 + q * 16 + col;
#pragma unroll 4
    for (int w0 = 0; w0 < N; w0 += 4) {
      const double2 a = mrow[(long long)w0 * 16], b = pcol[w0];
      p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, p1, 0, 0, 0);
      p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, p2, 0, 0, 0);
      p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x + a.y, b.x + b.y, p3, 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
    J.out[(lane >> 4) + 4 * i + 16 * col] = make_double2(p1[i] - p2[i], p3[i] - p1[i] - p2[i]);
}

// D = 24 / 32 (N a multiple of 4): the same product as 2 x 2 MFMA tiles of
// 16 x 16 per job (rows and columns past D are zero operands, not stored);
// lane l supplies M'[16 ta + l%16][k] and P2[k][16 tb + l%16] at
// k = bd*N + w0 + l/16.  One wavefront per job, 4 jobs per workgroup.
template <int D>
__global__ __launch_bounds__(256) void k_q2b_mfma(const Q2BJob *__restrict__ jobs, int njobs, int N, long long prow) {
  const int job = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (job >= njobs) return;
  const Q2BJob J = jobs[job];
  const int lane = threadIdx.x & 63, kk = lane >> 4, m = lane & 15;
  d4 acc[2][2][3];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[a][b][k] = d4{0.0, 0.0, 0.0, 0.0};
  const bool in1 = 16 + m < D;  // the second tile's row / column exists
  for (int bd = 0; bd < 4; ++bd) {
    const int g = c_gamma_row[J.gamma][bd];
    const double2 *p0 = J.peram2 + (long long)(J.b2 * D + m) * prow + (long long)(J.t * 4 + g) * N + kk;
    const double2 *p1 = p0 + 16LL * prow;
    const double2 *mr = J.Mp + ((long long)bd * N + kk) * D + m;
    for (int w0 = 0; w0 < N; w0 += 4) {
      const double2 zero = make_double2(0.0, 0.0);
      const double2 a0 = mr[(long long)w0 * D], a1 = in1 ? mr[(long long)w0 * D + 16] : zero;
      const double2 b0 = p0[w0], b1 = in1 ? p1[w0] : zero;
      const double2 av[2] = {a0, a1}, bv[2] = {b0, b1};
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
          acc[ta][tb][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ta].x, bv[tb].x, acc[ta][tb][0], 0, 0, 0);
          acc[ta][tb][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ta].y, bv[tb].y, acc[ta][tb][1], 0, 0, 0);
          acc[ta][tb][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ta].x + av[ta].y, bv[tb].x + bv[tb].y,
                                                                acc[ta][tb][2], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * ta + kk + 4 * i, col = 16 * tb + m;
        if (row < D && col < D) {
          const d4 &q1 = acc[ta][tb][0], &q2 = acc[ta][tb][1], &q3 = acc[ta][tb][2];
          J.out[(long long)col * D + row] = make_double2(q1[i] - q2[i], q3[i] - q1[i] - q2[i]);
        }
      }
}

__global__ void k_dagger(const double2 *__restrict__ in, double2 *__restrict__ out, int N, int count) {
  const long long total = (long long)N * N * count;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long mat = i / ((long long)N * N);
    const int r = int(i % N), c = int((i / N) % N);
    const double2 x = in[mat * N * N + (long long)r * N + c];  // in[c, r]
    out[i] = make_double2(x.x, -x.y);
  }
}

__global__ void k_zero2(double2 *p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = make_double2(0.0, 0.0);
}

__global__ void k_identity(double2 *out, int N, int count) {
  const long long total = (long long)N * N * count;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = int(i % N), c = int((i / N) % N);
    out[i] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// Batched D x D factor product (reference: operator*, DilutedFactor.cpp:6-51):
// one wavefront per output matrix, operands staged in LDS.
// ---------------------------------------------------------------------------
constexpr int kMaxD = 32;

template <typename T>
__global__ __launch_bounds__(64) void k_prod(const T *__restrict__ sX, const long long *__restrict__ Xt,
                                             const long long *__restrict__ Xq, const T *__restrict__ sY,
                                             const long long *__restrict__ Yt, const long long *__restrict__ Yq,
                                             T *__restrict__ out, long long obase, long long ostride,
                                             const ComboProd *__restrict__ combos, int D, int transposed) {
  __shared__ double2 A[kMaxD * kMaxD], B[kMaxD * kMaxD];
  const ComboProd c = combos[blockIdx.x];
  const int ti = blockIdx.y;
  const long long DD = (long long)D * D;
  const T *x = sX + (Xt[ti] + Xq[c.qn] + c.i) * DD;
  const T *y = sY + (Yt[ti] + Yq[c.qn] + c.j) * DD;
  for (int e = threadIdx.x; e < D * D; e += 64) {
    A[e] = ld2(x + e);
    B[e] = ld2(y + e);
  }
  __syncthreads();
  T *o = out + (obase + ti * ostride + c.o) * DD;
  for (int e = threadIdx.x; e < D * D; e += 64) {
    const int r = e % D, col = e / D;
    double2 s = make_double2(0.0, 0.0);
    for (int k = 0; k < D; ++k) cfma(s, A[k * D + r], B[col * D + k]);
    st2(o + (transposed ? col + r * D : e), s);
  }
}

// D = 16 factor products on FP64 MFMA: one 16 x 16 x 16 MFMA tile (4 k-steps x
// 3 real products, the Gauss form).  With the MFMA operand layout A[m = lane%16][k = lane/16],
// B[k = lane/16][n = lane%16] and result element i of lane l at
// (l/16 + 4i, l%16), lane l reads X at l + 64s (k-step s, coalesced) and Y at
// (4s + l/16) + 16 (l%16) (64-byte runs), and writes element l + 64i of the
// output (coalesced).  Computing C^T = Y^T X^T (Y as the A operand) makes that
// store C in column-major order; computing C = X Y (X as the A operand) makes
// it store C^T.  A wave computes kProdPerWave consecutive combinations of one
// time index; consecutive combinations that share the left matrix keep it in
// registers.
constexpr int kProdPerWave = 4;

template <typename T, bool kTransposed>
__global__ __launch_bounds__(256) void k_prod16(const T *__restrict__ sX, const long long *__restrict__ Xt,
                                                const long long *__restrict__ Xq, const T *__restrict__ sY,
                                                const long long *__restrict__ Yt, const long long *__restrict__ Yq,
                                                T *__restrict__ out, long long obase, long long ostride,
                                                const ComboProd *__restrict__ combos, int ncombo) {
  // XCD x (workgroups are dealt to the XCDs round-robin in dispatch order)
  // runs the x-th contiguous eighth of the (time, combination block)
  // sequence, so that the combinations that share operand matrices read them
  // through the same L2.
  const int NB = gridDim.x * gridDim.y, lin = blockIdx.y * gridDim.x + blockIdx.x;
  const int per = NB / 8, rem = NB % 8, xcd = lin % 8, jx = lin / 8;
  const int lb = xcd < rem ? xcd * (per + 1) + jx : rem * (per + 1) + (xcd - rem) * per + jx;
  const int ti = lb / gridDim.x, bx = lb % gridDim.x;
  // the wave index through readfirstlane: uniform, so that the combination
  // tables come through the scalar cache
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c0 = (bx * 4 + wave) * kProdPerWave;
  const int c1 = min(c0 + kProdPerWave, ncombo);
  if (c0 >= c1) return;
  const long long xt = Xt[ti], yt = Yt[ti];
  const int yoff = (lane >> 4) + 16 * (lane & 15);
  double2 xa[4];
  long long xcur = -1;
  // the right operand of the next combination is loaded while the current
  // one is multiplied and stored
  ComboProd cn = combos[c0];
  T yn[4];  // in the storage form, unpacked at use
  {
    const T *y = sY + (yt + Yq[cn.qn] + cn.j) * 256 + yoff;
#pragma unroll
    for (int s = 0; s < 4; ++s) yn[s] = y[4 * s];
  }
  for (int c = c0; c < c1; ++c) {
    const ComboProd cb = cn;
    double2 yb[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) yb[s] = ld2(&yn[s]);
    const long long xi = xt + Xq[cb.qn] + cb.i;
    if (xi != xcur) {
#pragma unroll
      for (int s = 0; s < 4; ++s) xa[s] = ld2(sX + xi * 256 + lane + 64 * s);
      xcur = xi;
    }
    if (c + 1 < c1) {
      cn = combos[c + 1];
      const T *y = sY + (yt + Yq[cn.qn] + cn.j) * 256 + yoff;
#pragma unroll
      for (int s = 0; s < 4; ++s) yn[s] = y[4 * s];
    }
    // C[r][c] = sum_k X[r][k] Y[k][c] from three real products:
    // P1 = Xr Yr, P2 = Xi Yi, P3 = (Xr + Xi)(Yr + Yi); re = P1 - P2, im = P3 - P1 - P2
    d4 p1 = d4{0.0, 0.0, 0.0, 0.0}, p2 = p1, p3 = p1;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double xs = xa[s].x + xa[s].y, ys = yb[s].x + yb[s].y;
      if (kTransposed) {
        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[s].x, yb[s].x, p1, 0, 0, 0);
        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[s].y, yb[s].y, p2, 0, 0, 0);
        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(xs, ys, p3, 0, 0, 0);
      } else {
        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(yb[s].x, xa[s].x, p1, 0, 0, 0);
        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(yb[s].y, xa[s].y, p2, 0, 0, 0);
        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ys, xs, p3, 0, 0, 0);
      }
    }
    T *o = out + (obase + ti * ostride + cb.o) * 256;
#pragma unroll
    for (int i = 0; i < 4; ++i) st2_nt(o + lane + 64 * i, make_double2(p1[i] - p2[i], p3[i] - p1[i] - p2[i]));
  }
}

// Summed products of the crossed four-factor trace (D = 16).  Q0 has one
// nonzero 4 x 4 Dirac block per block row and column, so a product with Q0
// on the left is sixteen 4 x 4 x 4 block products: four v_mfma_f64_4x4x4f64
// (four independent blocks each, the FP64 rate of the 16 x 16 x 4 form) per
// real product instead of four 16 x 16 x 4, three real products per complex
// one (Gauss).  Both kinds have Q0 on the left (kind 1 forms (Q2 Q0)^T =
// Q0^T Q2^T).  Lane l = 16 q + 4 g + r runs block slot g of every instruction
// (operand elements A[i][k] = lane i + 16 k + 4 g, B[k][j] = lane 16 k + j +
// 4 g, D[i][j] = lane 16 i + j + 4 g); instruction n computes output column
// block n from output row block g: C[g][n] = F[g][c] S[c][n] with F the Q0
// side (one nonzero block per row block, in column block c) and S the Q2 side.
// The lane's F element is the same for all n, so each lane loads one of the
// 64 nonzeros per term (a ring of kSpRing terms in flight), and four S
// elements from the unit's Q2 matrices in LDS (staged once per unit: every
// product of the unit reads them).  Instruction n's results are elements
// 64 n + (4 g + q + 16 r); before storing, lane l takes element 64 n + l from
// lane 16 (l & 3) + 4 ((l >> 2) & 3) + (l >> 4) (a swap of the 2-bit fields q
// and r), so that every store writes one KiB in lane order.  XCD x runs the
// x-th contiguous eighth of the (time, unit) sequence, so that the units of
// one time key read its Q0 slab through one L2.
constexpr int kSpRing = 4;
// LDS slot of element e = i + 16 j of a staged matrix: (i ^ j) + 16 j.  A
// wave's read of S[4 c + q][4 n + r] (kind 0) or S[4 n + r][4 c + q] (kind 1,
// transposed) then lands 16 consecutive lanes (fixed q, all g and r) on
// distinct 16-byte bank groups: i ^ j = 4 (c ^ n) + (q ^ r), with c running
// over all four blocks as g does (gamma permutations are XORs by a constant).
__device__ __forceinline__ int sp_swz(int e) { return e ^ ((e >> 4) & 15); }
template <typename T>
__global__ __launch_bounds__(64 * kSpWaves) void k_sprod16(
    const T *__restrict__ sQ0, const long long *__restrict__ Xt, const T *__restrict__ sQ2,
    const long long *__restrict__ Yt, T *__restrict__ out, long long obase, long long ostride,
    const STerm *__restrict__ terms, const SpUnit *__restrict__ units, int nunit) {
  __shared__ double2 S[kSpStage * 256];
  const int NB = gridDim.x * gridDim.y, lin = blockIdx.y * gridDim.x + blockIdx.x;
  const int per = NB / 8, rem = NB % 8, xcd = lin % 8, jx = lin / 8;
  const int lb = xcd < rem ? xcd * (per + 1) + jx : rem * (per + 1) + (xcd - rem) * per + jx;
  const int ti = lb / gridDim.x, u = lb % gridDim.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const SpUnit U = units[u];
  {
    const T *sb = sQ2 + (Yt[ti] + U.s0) * 256;
    for (int e = tid; e < U.np * 256; e += 64 * kSpWaves) S[(e & ~255) + sp_swz(e & 255)] = ld2(sb + e);
  }
  __syncthreads();
  const int t0 = U.tb[wave], t1 = U.tb[wave + 1];
  const T *q0 = sQ0 + Xt[ti] * 256;
  T *ob = out + (obase + ti * ostride) * 256;
  const int q = lane >> 4, g = (lane >> 2) & 3, r = lane & 3;
  // the column block c of F's nonzero block in row block g
  auto fblock = [&](int flags) {
    const int perm = flags >> 8;
    if (flags & 2) return (perm >> (2 * g)) & 3;  // F = Q0^T: Q0's block in column block g
    int c = 0;
#pragma unroll
    for (int d = 0; d < 4; ++d)
      if (((perm >> (2 * d)) & 3) == g) c = d;
    return c;
  };
  T fa[kSpRing];  // in the storage form, unpacked at use
  auto load = [&](int k, int t) {
    const STerm tm = terms[t];
    const int c = fblock(tm.flags);
    const int e = tm.flags & 2 ? (4 * c + q) + 16 * (4 * g + r) : (4 * g + r) + 16 * (4 * c + q);
    fa[k] = q0[(long long)tm.l * 256 + e];
  };
#pragma unroll
  for (int k = 0; k < kSpRing; ++k)
    if (t0 + k < t1) load(k, t0 + k);
  double p1[4], p2[4], p3[4];
  double2 sum[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    p1[n] = p2[n] = p3[n] = 0.0;
    sum[n] = make_double2(0.0, 0.0);
  }
  const int src = 4 * (16 * (lane & 3) + 4 * ((lane >> 2) & 3) + (lane >> 4));
  auto perm2 = [&](double2 v) {
    const int a = __builtin_amdgcn_ds_bpermute(src, __double2loint(v.x));
    const int b = __builtin_amdgcn_ds_bpermute(src, __double2hiint(v.x));
    const int c = __builtin_amdgcn_ds_bpermute(src, __double2loint(v.y));
    const int d = __builtin_amdgcn_ds_bpermute(src, __double2hiint(v.y));
    return make_double2(__hiloint2double(b, a), __hiloint2double(d, c));
  };
  for (int t = t0; t < t1; t += kSpRing) {
#pragma unroll
    for (int k = 0; k < kSpRing; ++k) {
      const int tt = t + k;
      if (tt < t1) {
        const STerm tm = terms[tt];
        const double2 fc = ld2(&fa[k]);
        if (tt + kSpRing < t1) load(k, tt + kSpRing);
        const bool k1 = tm.flags & 2, last = tm.flags & 1;
        const int c = fblock(tm.flags);
        const double2 *sm = S + tm.r * 256;
        double2 sc[4];
#pragma unroll
        for (int n = 0; n < 4; ++n)
          sc[n] = sm[sp_swz(k1 ? (4 * n + r) + 16 * (4 * c + q) : (4 * c + q) + 16 * (4 * n + r))];
        if (k1) {
#pragma unroll
          for (int n = 0; n < 4; ++n) p1[n] = p2[n] = p3[n] = 0.0;
        }
        const double fs = fc.x + fc.y;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          p1[n] = __builtin_amdgcn_mfma_f64_4x4x4f64(fc.x, sc[n].x, p1[n], 0, 0, 0);
          p2[n] = __builtin_amdgcn_mfma_f64_4x4x4f64(fc.y, sc[n].y, p2[n], 0, 0, 0);
          p3[n] = __builtin_amdgcn_mfma_f64_4x4x4f64(fs, sc[n].x + sc[n].y, p3[n], 0, 0, 0);
        }
        T *o = ob + (long long)tm.o * 256 + lane;
        if (k1) {
#pragma unroll
          for (int n = 0; n < 4; ++n) {
            const double2 v = make_double2(p1[n] - p2[n], p3[n] - p1[n] - p2[n]);
            st2_nt(o + 64 * n, perm2(v));
            sum[n] = cadd(sum[n], v);
          }
          if (last) {
#pragma unroll
            for (int n = 0; n < 4; ++n) {
              st2_nt(o + 256 + 64 * n, perm2(make_double2(-sum[n].y, sum[n].x)));
              sum[n] = make_double2(0.0, 0.0);
            }
          }
        } else if (last) {
#pragma unroll
          for (int n = 0; n < 4; ++n) {
            st2_nt(o + 64 * n, perm2(make_double2(p1[n] - p2[n], p3[n] - p1[n] - p2[n])));
            p1[n] = p2[n] = p3[n] = 0.0;
          }
        }
      }
    }
  }
}

// Traces of the summed-product tables as symmetric products per component of
// the trace graph (the trace ids of DilutedTraceFactory.cpp:161-190, each a sum
// of tr(X Y) = sum_{a,b} X[a,b] Y[b,a] over its matrix pairs, factor_to_trace
// in DilutedTrace.cpp:84-138).  One workgroup per (unit, slice pair): the
// unit's table matrices are read from memory once, stage by stage (two 8 x 8
// blocks per staged entry), and T = A B^T over the stage's K range runs on
// the 4 x 4 x 4 FP64 MFMA, one 4 x 4 tile of T (four row entries, four
// column entries) per instruction, its four blocks g splitting K, three real
// products per complex one (Gauss).  Lane l = i + 4 g + 16 k supplies entry i
// of the tile's rows (A) and of its columns (B) at K element (row 2 s +
// (k >> 1), column 4 (k & 1) + g) of the stage block for sub-step s: A reads
// that element of its slot, B the transposed element of the paired slot.
// LDS position of element (r, c) of staged position p: 64 p + ((8 r + (c ^ r))
// ^ 4 (p & 3)); the 16 lanes of one k (four entries, four consecutive
// elements of a block row for A, of a block column for B) then fall on
// sixteen distinct 16-byte bank groups.  Two stage buffers: while stage k
// is multiplied, stage k + 1 waits in registers for its buffer and stage
// k + 2's loads are in flight (two register sets); one barrier per stage.
// Lanes 16 i + j hold T[i][j] of a tile after the sum over the blocks g (row
// rotations by 8 and 4).
constexpr int kSymPass = (2 * kSymEntries + kSymWaves - 1) / kSymWaves;  // staged blocks per wave and stage
__device__ __forceinline__ int sym_pos(int p, int r, int c) { return 64 * p + ((8 * r + (c ^ r)) ^ ((p & 3) << 2)); }
template <typename T>
__global__ __launch_bounds__(64 * kSymWaves) __attribute__((amdgpu_waves_per_eu(kSymWaves / 4))) void k_trace_sym(
    const T *__restrict__ slab, const long long *__restrict__ Tt, const SymUnit *__restrict__ units, int ntime,
    const int *__restrict__ stages, const int *__restrict__ tiles, double2 *__restrict__ part, long long pstride) {
  // two stage buffers (stage k in buffer k % 2), two 8 x 8 blocks per position
  __shared__ double2 Ls[2][2][kSymEntries * 64];
  const int u = blockIdx.x / ntime, ti = blockIdx.x % ntime;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // unit fields, table entries and tile headers are wave-uniform: through
  // readfirstlane the branches on them are scalar
  auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  const SymUnit *U = units + u;
  const int ns = uni(U->ns), nst = uni(U->nst), rowlen = 1 + 2 * ns;
  const int *st = stages + uni(U->st_off);
  // staging: wave w loads blocks b = w + kSymWaves q (position b / 2, slot
  // b % 2), lane (r, c) = (lane / 8, lane % 8) of the 8 x 8 block
  const int cr = lane >> 3, cc = lane & 7;
  const T *tb = slab + Tt[ti] * 256 + cr * 16 + cc;
  // stage k's loads go to register set k % 2: set (k + 1) % 2 holds stage
  // k + 1 (issued one stage earlier) while stage k is multiplied.  All
  // offsets of a stage are read first (scalar loads), then its block loads
  // go out back to back.  Every wave issues the same kSymPass loads per
  // stage, whatever the unit (blocks past it and zero blocks, offset -1,
  // load block 0 and are not used: bit q of the set's mask), and the stages
  // past the last reload the last one, so that the compiler's count of
  // loads in flight is exact: the write of stage k + 1 to LDS then waits
  // only for its own loads, not for stage k + 2's.
  T pre0[kSymPass], pre1[kSymPass];
  unsigned zero0 = 0, zero1 = 0;
  const int nb = 2 * ns;
  auto fetch = [&](int k, T(&pre)[kSymPass], unsigned &zero) {
    const int *row = st + min(k, nst - 1) * rowlen + 1;
    int off[kSymPass];
#pragma unroll
    for (int q = 0; q < kSymPass; ++q) off[q] = kSymWaves * q + wave < nb ? uni(row[kSymWaves * q + wave]) : -1;
    zero = 0;
#pragma unroll
    for (int q = 0; q < kSymPass; ++q) {
      zero |= off[q] < 0 ? 1u << q : 0u;
      pre[q] = tb[max(off[q], 0)];
    }
  };
  auto commit = [&](T(&pre)[kSymPass], unsigned zero, int buf) {
#pragma unroll
    for (int q = 0; q < kSymPass; ++q) {
      const int b = kSymWaves * q + wave;
      if (b < nb) Ls[buf][b & 1][sym_pos(b >> 1, cr, cc)] = (zero >> q) & 1 ? make_double2(0.0, 0.0) : ld2(&pre[q]);
    }
  };
  // this wave's tiles
  const int t0 = uni(U->wt[wave]), nt = uni(U->wt[wave + 1]) - t0;
  const int *tl = tiles + 17LL * (uni(U->tile_off) + t0);
  int th[kSymWaveTiles];
#pragma unroll
  for (int t = 0; t < kSymWaveTiles; ++t) th[t] = t < nt ? uni(tl[17 * t]) : 0;
  double p1[kSymWaveTiles], p2[kSymWaveTiles], p3[kSymWaveTiles];
#pragma unroll
  for (int t = 0; t < kSymWaveTiles; ++t) p1[t] = p2[t] = p3[t] = 0.0;
  // per sub-step s: the lane's A and B element within a group of four positions
  const int li = lane & 3, lg = (lane >> 2) & 3, lk = lane >> 4;
  int offA[4], offB[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int r = 2 * s + (lk >> 1), c = 4 * (lk & 1) + lg;
    offA[s] = sym_pos(li, r, c);
    offB[s] = sym_pos(li, c, r);
  }
  auto compute = [&](int buf, bool self) {
    int curR = -1;
    double2 a[2][4];
#pragma unroll
    for (int t = 0; t < kSymWaveTiles; ++t) {
      if (t < nt) {
        const int R = th[t] & 255, C = th[t] >> 8;
        if (R != curR) {
          curR = R;
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int s = 0; s < 4; ++s) a[h][s] = Ls[buf][h][256 * R + offA[s]];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const double2 *bp = Ls[buf][self ? h : 1 - h] + 256 * C;
          double2 b[4];
#pragma unroll
          for (int s = 0; s < 4; ++s) b[s] = bp[offB[s]];
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const double2 x = a[h][s], y = b[s];
            p1[t] = __builtin_amdgcn_mfma_f64_4x4x4f64(x.x, y.x, p1[t], 0, 0, 0);
            p2[t] = __builtin_amdgcn_mfma_f64_4x4x4f64(x.y, y.y, p2[t], 0, 0, 0);
            p3[t] = __builtin_amdgcn_mfma_f64_4x4x4f64(x.x + x.y, y.x + y.y, p3[t], 0, 0, 0);
          }
        }
      }
    }
  };
  // one stage: the loads of stage k + 2 go out into the set stage k left,
  // stage k is multiplied, stage k + 1 is written to the other buffer (whose
  // stage k - 1 every wave finished before the last barrier)
  // (nst is even; the last step's write goes to a buffer nobody reads)
  auto step = [&](int k, T(&cur)[kSymPass], unsigned &zc, T(&nxt)[kSymPass], unsigned zn, int buf) {
    const int flags = uni(st[k * rowlen]);
    fetch(k + 2, cur, zc);
    compute(buf, flags & 1);
    commit(nxt, zn, buf ^ 1);
    __syncthreads();
  };
  fetch(0, pre0, zero0);
  fetch(1, pre1, zero1);
  commit(pre0, zero0, 0);
  __syncthreads();
  for (int k = 0; k < nst; k += 2) {
    step(k, pre0, zero0, pre1, zero1, 0);
    step(k + 1, pre1, zero1, pre0, zero0, 1);
  }
  double2 *pb = part + (long long)ti * pstride;
#pragma unroll
  for (int t = 0; t < kSymWaveTiles; ++t) {
    if (t < nt) {
      double re = p1[t] - p2[t], im = p3[t] - p1[t] - p2[t];
      re += dpp_d<0x128>(re);  // row_ror:8
      im += dpp_d<0x128>(im);
      re += dpp_d<0x124>(re);  // row_ror:4
      im += dpp_d<0x124>(im);
      if (lg == 0) {
        const int slot = tl[17 * t + 1 + 4 * lk + li];
        if (slot >= 0) pb[slot] = make_double2(re, im);
      }
    }
  }
}

// D = 24 and D = 32 factor products on FP64 MFMA: one wavefront per
// product, the D x D output as 2 x 2 tiles of 16 x 16 (rows and columns past
// D of the D = 24 case are zero operands and are not stored), D / 4 k-steps of
// 16x16x4 MFMAs, three real products per complex one (Gauss).  As in k_prod16
// the stored layout decides the operand roles: C (column-major) is computed
// as C^T = Y^T X^T with Y as the A operand, so that a lane's results are 16
// consecutive rows of one column; C^T (kTransposed) as C = X Y.
template <typename T, int D, bool kTransposed>
__global__ __launch_bounds__(256) void k_prod_mfma(const T *__restrict__ sX, const long long *__restrict__ Xt,
                                                   const long long *__restrict__ Xq, const T *__restrict__ sY,
                                                   const long long *__restrict__ Yt, const long long *__restrict__ Yq,
                                                   T *__restrict__ out, long long obase, long long ostride,
                                                   const ComboProd *__restrict__ combos, int ncombo) {
  static_assert(D == 24 || D == 32, "k_prod_mfma is for D = 24 and 32");
  constexpr int KS = D / 4;  // k-steps of 4
  // XCD-contiguous mapping of the (time, combination block) sequence, as k_prod16
  const int NB = gridDim.x * gridDim.y, lin = blockIdx.y * gridDim.x + blockIdx.x;
  const int per = NB / 8, rem = NB % 8, xcd = lin % 8, jx = lin / 8;
  const int lb = xcd < rem ? xcd * (per + 1) + jx : rem * (per + 1) + (xcd - rem) * per + jx;
  const int ti = lb / gridDim.x, bx = lb % gridDim.x;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = bx * 4 + wave;
  if (c >= ncombo) return;
  const ComboProd cb = combos[c];
  const T *x = sX + (Xt[ti] + Xq[cb.qn] + cb.i) * (long long)(D * D);
  const T *y = sY + (Yt[ti] + Yq[cb.qn] + cb.j) * (long long)(D * D);
  const int m = lane & 15, kk = lane >> 4;
  // A[t][s]: element (16 t + m, 4 s + kk) of the A matrix, B[t][s]: element
  // (4 s + kk, 16 t + m) of the B matrix, zero past D
  double2 A[2][KS], B[2][KS];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int r = 16 * t + m, k = 4 * s + kk;
      const bool in = r < D;
      if (kTransposed) {  // A = X (r, k), B = Y (k, r)
        A[t][s] = in ? ld2(x + k * D + r) : make_double2(0.0, 0.0);
        B[t][s] = in ? ld2(y + r * D + k) : make_double2(0.0, 0.0);
      } else {  // A = Y^T (r, k) = Y (k, r), B = X^T (k, r) = X (r, k)
        A[t][s] = in ? ld2(y + r * D + k) : make_double2(0.0, 0.0);
        B[t][s] = in ? ld2(x + k * D + r) : make_double2(0.0, 0.0);
      }
    }
  T *o = out + (obase + ti * ostride + cb.o) * (long long)(D * D);
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb) {
      d4 p1 = d4{0.0, 0.0, 0.0, 0.0}, p2 = p1, p3 = p1;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const double2 a = A[ta][s], b = B[tb][s];
        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, p1, 0, 0, 0);
        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, p2, 0, 0, 0);
        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x + a.y, b.x + b.y, p3, 0, 0, 0);
      }
      // result element i of lane l: row 16 ta + l/16 + 4 i, column 16 tb + l%16
      // of A B; stored at (column 16 ta + ..) x D + row 16 tb + l%16 of the
      // output (C when A = Y^T, C^T when A = X), i.e. 16 consecutive elements
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ar = 16 * ta + kk + 4 * i, bc = 16 * tb + m;
        if (ar < D && bc < D) st2(o + (long long)ar * D + bc, make_double2(p1[i] - p2[i], p3[i] - p1[i] - p2[i]));
      }
    }
}

// ---------------------------------------------------------------------------
// Non-separable traces tr(X Y) = sum_{a,b} X[a,b] Y[b,a] (reference:
// factor_to_trace(L, R), DilutedTrace.cpp:84-138), left-stationary: one
// workgroup per (time, block of slots).  The block's left matrices (any list
// of operator entries and combinations) are staged chunk by chunk in LDS,
// transposed, so that element e of a right matrix meets its partner at the
// same index.  Each wavefront streams the right matrices of its tasks into
// registers once (EPL = D*D/64 elements per lane; up to PF tasks ahead are in
// flight, since the kernel is bound by the latency of these loads) and
// multiplies them with one or two staged left matrices.  At the end of a run
// (or of the wave's task range) the wave sum is added to the wave's own copy
// of the run's slot; the copies are summed in wave order at the end, so the
// summation order is fixed by the plan.
// ---------------------------------------------------------------------------
constexpr int kLsStageBytes = 52 * 1024;
// LDS slot of element e of a staged matrix: the low four bits of the 16-byte
// slot index are XORed with the next four, so that both the lane-contiguous
// reads and the transposing writes (stride D slots) hit distinct banks.
__device__ __forceinline__ int ls_swz(int e) { return e ^ ((e >> 4) & 15); }

template <int D, typename TS>
__global__ __launch_bounds__(64 * kLsWaves) __attribute__((amdgpu_waves_per_eu(3))) void k_trace_lstat(
    const TS *__restrict__ sX, const long long *__restrict__ Xt, const TS *__restrict__ sY,
    const long long *__restrict__ Yt, const LsBlock *__restrict__ blocks, int nblock, const LsChunk *__restrict__ chunks,
    const int *__restrict__ left, const LsTask *__restrict__ tasks, const long long *__restrict__ slot_dst,
    double2 *__restrict__ out, long long obase, long long ostride, int rt) {
  constexpr int DD = D * D, EPL = DD / 64, CL = kLsStageBytes / (DD * 16);
  constexpr int PF = EPL <= 4 ? 4 : (EPL <= 8 ? 2 : 1);  // right matrices in flight per wave
  __shared__ double2 Ls[CL * DD];
  __shared__ double2 acc[kLsWaves][kLsSlots];
  // Workgroups are dealt round-robin to the 8 XCDs, each with its own L2:
  // map them so that XCD x runs the x-th contiguous eighth of the (time,
  // block) sequence, where neighbouring blocks share right matrices.
  const int G = gridDim.x, per = G / 8, rem = G % 8, x = blockIdx.x % 8, j = blockIdx.x / 8;
  const int logical = x < rem ? x * (per + 1) + j : rem * (per + 1) + (x - rem) * per + j;
  const int ti = logical / nblock;
  const LsBlock B = blocks[logical % nblock];
  // the wave index through readfirstlane: uniform, so that per-wave table
  // entries are read as scalars
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int k = tid; k < kLsWaves * kLsSlots; k += 64 * kLsWaves) (&acc[0][0])[k] = make_double2(0.0, 0.0);
  const TS *xb = sX + Xt[ti] * DD;
  const TS *yb = sY + Yt[ti] * DD;
  // The next chunk's left matrices are loaded into registers while the
  // current chunk's tasks run, and written to LDS between the two.
  constexpr int SPT = (CL * DD + 64 * kLsWaves - 1) / (64 * kLsWaves);
  TS pre[SPT];
  auto fetch = [&](int c) {
    const LsChunk C = chunks[B.chunk_off + c];
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      const int idx = tid + 64 * kLsWaves * k;
      // D*D is a multiple of 64, so the staged matrix m is uniform in a wave
      // and its offset comes through the scalar cache.
      const int m = __builtin_amdgcn_readfirstlane(idx / DD);
      if (m < C.n) pre[k] = xb[(long long)left[C.loff + m] * DD + idx % DD];
    }
  };
  auto commit = [&](int c) {
    const int n = chunks[B.chunk_off + c].n;
    __syncthreads();  // the previous chunk is no longer read
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      const int idx = tid + 64 * kLsWaves * k, m = idx / DD, f = idx % DD;
      if (idx < n * DD) Ls[m * DD + ls_swz(rt ? f : f / D + (f % D) * D)] = ld2(&pre[k]);
    }
    __syncthreads();
  };
  int so[EPL];  // swizzled LDS slots of this lane's elements lane + 64 k
#pragma unroll
  for (int k = 0; k < EPL; ++k) so[k] = ls_swz(lane + 64 * k);
  // Each wave's tasks of the block form one contiguous stream, chunk by
  // chunk, every chunk's share padded to a multiple of PF (null tasks have
  // ia < 0).  The ring of PF right matrices in flight runs through the whole
  // stream: the first tasks of chunk c + 1 are loaded while chunk c finishes,
  // so only the LDS staging waits at a chunk boundary.  The descriptors sit
  // one per lane, 64 stream positions at a time; a new batch starts before a
  // round would read past it (a round reads up to 2 PF - 1 positions ahead).
  const int s0 = chunks[B.chunk_off].wbeg[wave];
  const int ns = chunks[B.chunk_off + B.nchunk - 1].wbeg[wave] + chunks[B.chunk_off + B.nchunk - 1].wcnt[wave] - s0;
  int bb = 0;  // stream position of the descriptor batch
  int qy = 0, qab = -1, qs = -1;
  auto batch = [&]() {
    qy = 0;
    qab = -1;
    qs = -1;
    if (bb + lane < ns) {
      const LsTask q = tasks[s0 + bb + lane];
      qy = q.yrel;
      qab = int((unsigned short)q.ia) | (int(q.ib) << 16);
      qs = q.slot;
    }
  };
  fetch(0);
  // ring of right matrices in the storage form, unpacked at use (an unpack
  // right after the load would wait for it there)
  TS r[PF][EPL];
  if (ns > 0) {
    batch();
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const long long y = __builtin_amdgcn_readlane(qy, min(u, ns - 1));
#pragma unroll
      for (int k = 0; k < EPL; ++k) r[u][k] = yb[y * DD + lane + 64 * k];
    }
  }
  double re = 0.0, im = 0.0;
  int pos = 0;  // stream position of the next task
  for (int c = 0; c < B.nchunk; ++c) {
    const int cnt = chunks[B.chunk_off + c].wcnt[wave];
    commit(c);
    if (c + 1 < B.nchunk) fetch(c + 1);
    for (int e = pos + cnt; pos < e; pos += PF) {
      if (pos - bb > 64 - 2 * PF) {  // a round reads up to pos - bb + 2 PF - 1
        bb = pos;
        batch();
      }
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int ab = __builtin_amdgcn_readlane(qab, pos + u - bb);
        const int ia = (short)(ab & 0xffff), ib = ab >> 16;
        if (ia >= 0) {
          // the partners are read together and added first:
          // tr(L_a R) + tr(L_b R) = sum_e (L_a + L_b)[e] R[e]
          const double2 *la = Ls + ia * DD;
          double2 w[EPL];
#pragma unroll
          for (int k = 0; k < EPL; ++k) w[k] = la[so[k]];
          if (ib >= 0) {
            const double2 *lb = Ls + ib * DD;
#pragma unroll
            for (int k = 0; k < EPL; ++k) {
              const double2 b = lb[so[k]];
              w[k].x += b.x;
              w[k].y += b.y;
            }
          }
#pragma unroll
          for (int k = 0; k < EPL; ++k) {
            const double2 v = ld2(&r[u][k]);
            re = fma(w[k].x, v.x, re);
            re = fma(-w[k].y, v.y, re);
            im = fma(w[k].x, v.y, im);
            im = fma(w[k].y, v.x, im);
          }
          const int slot = __builtin_amdgcn_readlane(qs, pos + u - bb);
          if (slot >= 0) {
            re = wave_sum(re);
            im = wave_sum(im);
            if (lane == 0) acc[wave][slot] = cadd(acc[wave][slot], make_double2(re, im));
            re = 0.0;
            im = 0.0;
          }
        }
        // refill this ring slot with the task PF positions ahead (past the
        // stream's end: its last task again, so that every load is issued)
        const long long y = __builtin_amdgcn_readlane(qy, min(pos + u + PF, ns - 1) - bb);
#pragma unroll
        for (int k = 0; k < EPL; ++k) r[u][k] = yb[y * DD + lane + 64 * k];
      }
    }
  }
  __syncthreads();
  for (int k = tid; k < B.nslot; k += 64 * kLsWaves) {
    double2 v = acc[0][k];
#pragma unroll
    for (int w = 1; w < kLsWaves; ++w) v = cadd(v, acc[w][k]);
    out[obase + ti * ostride + slot_dst[B.slot_off + k]] = v;
  }
}

// Queued form of the same traces: one workgroup per resident slot.  The
// units (time, block of one chunk) of XCD x's contiguous eighth of the unit
// sequence are claimed one at a time from counter x (then from the other
// counters once x's range is spent), so that the units running together on
// one XCD are neighbours in plan order and find each other's right matrices
// in that XCD's L2.  Claims run two units ahead: the ring of right matrices
// runs on from unit g into unit g+1; at the start of unit g the workgroup
// loads unit g+1's first task batch and unit g+2's record, and at its end
// unit g+1's left matrices (into registers).  A unit record is kLsUnitHdr ints (n, slot_off, nslot,
// wbeg[kLsWaves], wcnt[kLsWaves], 0), each wave's first PF right matrices
// (0 past the wave's range), then the unit's left matrix offsets, so that no
// descriptor load waits on another.  The task array ends with one
// null task, so that an empty range still has a readable first entry.
// Results do not depend on which workgroup runs a unit: each unit writes its
// own partial sums.
template <int D, typename TS>
__global__ __launch_bounds__(64 * kLsWaves) __attribute__((amdgpu_waves_per_eu(2))) void k_trace_lstat_q(
    const TS *__restrict__ sX, const long long *__restrict__ Xt, const TS *__restrict__ sY,
    const long long *__restrict__ Yt, const int *__restrict__ units, int nunit, const LsTask *__restrict__ tasks,
    const long long *__restrict__ slot_dst, double2 *__restrict__ out, long long obase, long long ostride, int ntime,
    int rt, int *__restrict__ qcnt, int win) {
  constexpr int DD = D * D, EPL = DD / 64;
  constexpr int PF = EPL <= 4 ? 6 : (EPL <= 8 ? 2 : 1);  // trace_lstat_qring(D)
  constexpr int YR = kLsUnitHdr + kLsWaves * PF;  // offset of the staged left matrices in a record
  constexpr int CL = kLsStageBytes / (DD * 16) < 64 - YR ? kLsStageBytes / (DD * 16) : 64 - YR;
  constexpr int UR = YR + CL;  // ints per unit record, one per lane
  static_assert(UR <= 64, "a unit record must fit one int per lane");
  static_assert(kLsSlots <= 64 * kLsWaves, "one thread per slot");
  __shared__ double2 Ls[CL * DD];
  __shared__ double2 acc[kLsWaves][kLsSlots];
  __shared__ int qslot[2];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int total = nunit * ntime, per = total / 8, rem = total % 8, home = blockIdx.x % 8;
  // Thread 0's claims: a fetch-and-add on counter (home + tried) % 8, whose
  // value is a position in that counter's range; a spent range moves the
  // claim on to the next counter (synchronously: at most eight times).
  int tried = 0;
  auto resolve = [&](int v) -> int {
    for (;;) {
      const int x = (home + tried) & 7;
      if (v < per + (x < rem ? 1 : 0)) return x * per + min(x, rem) + v;
      if (++tried == 8) return -1;
      v = atomicAdd(&qcnt[(home + tried) & 7], 1);
    }
  };
  for (int k = tid; k < kLsWaves * kLsSlots; k += 64 * kLsWaves) (&acc[0][0])[k] = make_double2(0.0, 0.0);
  int qv = 0;
  if (tid == 0) {
    const int g0 = resolve(atomicAdd(&qcnt[home], 1));
    const int g1 = g0 >= 0 ? resolve(atomicAdd(&qcnt[(home + tried) & 7], 1)) : -1;
    if (g1 >= 0 && tried < 8) qv = atomicAdd(&qcnt[(home + tried) & 7], 1);
    qslot[0] = g0;
    qslot[1] = g1;
  }
  __syncthreads();
  int g = __builtin_amdgcn_readfirstlane(qslot[0]), gn = __builtin_amdgcn_readfirstlane(qslot[1]);
  if (g < 0) return;
  // unit record (one int per lane), and the unit's time bases
  auto rec = [&](int u) { return units[(long long)(u % nunit) * UR + min(lane, UR - 1)]; };
  auto tbase = [&](const long long *tb, int u) { return tb[u / nunit]; };
  int ud = rec(g), udn = rec(max(gn, 0));
  long long xt = tbase(Xt, g), yt = tbase(Yt, g), xtn = tbase(Xt, max(gn, 0)), ytn = tbase(Yt, max(gn, 0));
  constexpr int SPT = (CL * DD + 64 * kLsWaves - 1) / (64 * kLsWaves);
  TS pre[SPT];
  // a wave-uniform 64-bit value through the scalar unit
  auto uni = [](long long v) {
    return (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(int(v >> 32)) << 32) |
                       (unsigned)__builtin_amdgcn_readfirstlane(int(v)));
  };
  auto fetch = [&](int r, long long xbase) {
    const TS *xb = sX + uni(xbase) * DD;
    const int n = __builtin_amdgcn_readlane(r, 0);
    if constexpr (DD == 64 * kLsWaves) {  // element tid of matrix k
#pragma unroll
      for (int k = 0; k < SPT; ++k)
        if (k < n) pre[k] = xb[(long long)__builtin_amdgcn_readlane(r, YR + k) * DD + tid];
    } else {
#pragma unroll
      for (int k = 0; k < SPT; ++k) {
        const int idx = tid + 64 * kLsWaves * k;
        const int m = __builtin_amdgcn_readfirstlane(idx / DD);
        if (m < n) pre[k] = xb[(long long)__builtin_amdgcn_readlane(r, YR + m) * DD + idx % DD];
      }
    }
  };
  auto commit = [&](int n) {
    if constexpr (DD == 64 * kLsWaves) {
      // element tid of matrix k: one LDS address, the matrices at constant offsets
      double2 *dst = Ls + ls_swz(rt ? tid : tid / D + (tid % D) * D);
#pragma unroll
      for (int k = 0; k < SPT; ++k)
        if (k < n) dst[k * DD] = ld2(&pre[k]);
    } else {
#pragma unroll
      for (int k = 0; k < SPT; ++k) {
        const int idx = tid + 64 * kLsWaves * k, m = idx / DD, f = idx % DD;
        if (idx < n * DD) Ls[m * DD + ls_swz(rt ? f : f / D + (f % D) * D)] = ld2(&pre[k]);
      }
    }
  };
  // one task batch: positions b0 + lane of this wave's range of a unit
  // (lanes past the range repeat its last task; those positions are never read)
  auto batch = [&](int r, int b0, int &qy, int &qab, int &qs) {
    const int s0 = __builtin_amdgcn_readlane(r, 3 + wave), ns = __builtin_amdgcn_readlane(r, 3 + kLsWaves + wave);
    const int *t = reinterpret_cast<const int *>(tasks + s0 + min(b0 + lane, max(ns - 1, 0)));
    qy = t[0];
    qab = t[1];
    qs = t[2];
  };
  auto ybase = [&](long long yoff) { return sY + uni(yoff) * DD; };
  // an empty use: the loads that produced these registers are waited for here
  auto keep3 = [](int &a, int &b, int &c) { asm volatile("" : "+v"(a), "+v"(b), "+v"(c)); };
  int so[EPL];  // swizzled LDS slots of this lane's elements lane + 64 k
#pragma unroll
  for (int k = 0; k < EPL; ++k) so[k] = ls_swz(lane + 64 * k);
  fetch(ud, xt);
  int qy, qab, qs;
  batch(ud, 0, qy, qab, qs);
  // ring of right matrices in the storage form, unpacked at use (an unpack
  // right after the load would wait for it there)
  TS r[PF][EPL];
  {
    const TS *yb = ybase(yt);
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const long long y = __builtin_amdgcn_readlane(qy, u);
#pragma unroll
      for (int k = 0; k < EPL; ++k) r[u][k] = yb[y * DD + lane + 64 * k];
    }
  }
  int pti = 0, pso = 0, pns = 0;  // the previous unit's slots, summed at the next unit's start
  for (;;) {
    __syncthreads();  // every wave is done with the previous unit
    if (tid < pns) {
      double2 v = acc[0][tid];
      acc[0][tid] = make_double2(0.0, 0.0);
#pragma unroll
      for (int w = 1; w < kLsWaves; ++w) {
        v = cadd(v, acc[w][tid]);
        acc[w][tid] = make_double2(0.0, 0.0);
      }
      out[obase + pti * ostride + pso + tid] = v;
    }
    commit(__builtin_amdgcn_readlane(ud, 0));
    // this unit's first batch and the next unit's record were loaded before
    // this unit's left matrices, on which commit has just waited
    keep3(qy, qab, qs);
    keep3(ud, udn, udn);
    if (tid == 0) qslot[0] = gn >= 0 && tried < 8 ? resolve(qv) : -1;
    __syncthreads();
    const int g2 = __builtin_amdgcn_readfirstlane(qslot[0]);
    const int nsn = gn >= 0 ? __builtin_amdgcn_readlane(udn, 3 + kLsWaves + wave) : 0;
    const TS *yb = ybase(yt), *ybn = ybase(ytn);
    // The record and time bases of the unit after next and the next unit's
    // first task batch (small loads that hit L2) go out now; its left
    // matrices (long-latency loads) only behind this unit's last ring refill,
    // since vector loads complete in issue order and each ring wait would
    // otherwise wait for them too.
    const int ud2 = rec(max(g2, 0));
    const long long xt2 = tbase(Xt, max(g2, 0)), yt2 = tbase(Yt, max(g2, 0));
    int qyN = 0, qabN = -1, qsN = -1;
    if (gn >= 0) batch(udn, 0, qyN, qabN, qsN);
    const int ns = __builtin_amdgcn_readlane(ud, 3 + kLsWaves + wave);
    double re = 0.0, im = 0.0;
    // Windows of win stream positions share one descriptor batch: a round
    // reads up to pos - bb + 2 PF - 1 <= 63, and a window is a whole number
    // of rounds (win a multiple of PF, at most 64 - PF; the launcher checks).
    // Inside a window the only loads are the ring refills, so that their
    // waits stay exact.
    for (int bb = 0; bb < ns; bb += win) {
      if (bb > 0) {
        batch(ud, bb, qy, qab, qs);
        keep3(qy, qab, qs);
      }
      const int wend = min(ns, bb + win);
      for (int pos = bb; pos < wend; pos += PF) {
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int ab = __builtin_amdgcn_readlane(qab, pos + u - bb);
        const int ia = (short)(ab & 0xffff), ib = ab >> 16;
        if (ia >= 0) {
          // tr(L_a R) + tr(L_b R) = sum_e (L_a + L_b)[e] R[e]
          const double2 *la = Ls + ia * DD;
          double2 w[EPL];
#pragma unroll
          for (int k = 0; k < EPL; ++k) w[k] = la[so[k]];
          if (ib >= 0) {
            const double2 *lb = Ls + ib * DD;
#pragma unroll
            for (int k = 0; k < EPL; ++k) {
              const double2 b = lb[so[k]];
              w[k].x += b.x;
              w[k].y += b.y;
            }
          }
#pragma unroll
          for (int k = 0; k < EPL; ++k) {
            const double2 v = ld2(&r[u][k]);
            re = fma(w[k].x, v.x, re);
            re = fma(-w[k].y, v.y, re);
            im = fma(w[k].x, v.y, im);
            im = fma(w[k].y, v.x, im);
          }
          const int slot = __builtin_amdgcn_readlane(qs, pos + u - bb);
          if (slot >= 0) {
            re = wave_sum(re);
            im = wave_sum(im);
            if (lane == 0) acc[wave][slot] = cadd(acc[wave][slot], make_double2(re, im));
            re = 0.0;
            im = 0.0;
          }
        }
        // refill this ring slot with the stream position PF ahead: this
        // unit's task, else the next unit's, else (past the end) matrix 0
        const int p = pos + u + PF;
        const TS *base = p < ns ? yb : (p - ns < nsn ? ybn : yb);
        const long long y =
            p < ns ? __builtin_amdgcn_readlane(qy, p - bb)
                   : (p - ns < nsn ? __builtin_amdgcn_readlane(udn, kLsUnitHdr + wave * PF + min(p - ns, PF - 1)) : 0);
#pragma unroll
        for (int k = 0; k < EPL; ++k) r[u][k] = base[y * DD + lane + 64 * k];
      }
      }
    }
    if (ns == 0) {  // no task of this wave in the unit: the ring starts the next unit
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const TS *base = u < nsn ? ybn : yb;
        const long long y = u < nsn ? __builtin_amdgcn_readlane(udn, kLsUnitHdr + wave * PF + u) : 0;
#pragma unroll
        for (int k = 0; k < EPL; ++k) r[u][k] = base[y * DD + lane + 64 * k];
      }
    }
    // the claim after next, then the next unit's left matrices
    if (tid == 0 && g2 >= 0 && tried < 8) qv = atomicAdd(&qcnt[(home + tried) & 7], 1);
    if (gn >= 0) fetch(udn, xtn);
    pti = g / nunit;
    pso = __builtin_amdgcn_readlane(ud, 1);
    pns = __builtin_amdgcn_readlane(ud, 2);
    if (gn < 0) break;
    g = gn;
    ud = udn;
    xt = xtn;
    yt = ytn;
    qy = qyN;
    qab = qabN;
    qs = qsN;
    gn = g2;
    udn = ud2;
    xtn = xt2;
    ytn = yt2;
  }
  __syncthreads();
  if (tid < pns) {
    double2 v = acc[0][tid];
#pragma unroll
    for (int w = 1; w < kLsWaves; ++w) v = cadd(v, acc[w][tid]);
    out[obase + pti * ostride + pso + tid] = v;
  }
}

__global__ void k_ls_qreset(int *q) {
  if (threadIdx.x < 8) q[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// Traces fused with their mask-group sums (reference: factor_to_trace,
// DilutedTrace.cpp:84-165, followed by the sums inside inner_product /
// resolve_request1): one wavefront per (time, trace id, mask group); each lane
// accumulates its share of sum_combos sum_{a,b} X[a,b] Y[b,a] over all
// combinations of the group, then one wave reduction.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_trace_groups(
    const T *__restrict__ sX, const long long *__restrict__ Xt, const long long *__restrict__ Xq,
    const T *__restrict__ sY, const long long *__restrict__ Yt, const long long *__restrict__ Yq,
    const GroupItem *__restrict__ items, const int2 *__restrict__ pairs, double2 *__restrict__ out, long long obase,
    long long ostride, int D, int rt) {
  const GroupItem g = items[blockIdx.x];
  const int ti = blockIdx.y;
  const long long DD = (long long)D * D;
  const T *xb = sX + (Xt[ti] + Xq[g.qn]) * DD;
  double re = 0.0, im = 0.0;
  if (sY) {
    const T *yb = sY + (Yt[ti] + Yq[g.qn]) * DD;
    for (int c = 0; c < g.n; ++c) {
      const int2 p = pairs[g.list_off + c];
      const T *x = xb + p.x * DD;
      const T *y = yb + p.y * DD;
      for (int e = threadIdx.x; e < D * D; e += 64) {
        const int a = e % D, b = e / D;
        const double2 u = ld2(x + e), v = ld2(y + (rt ? e : a * D + b));
        re = fma(u.x, v.x, re);
        re = fma(-u.y, v.y, re);
        im = fma(u.x, v.y, im);
        im = fma(u.y, v.x, im);
      }
    }
  } else {
    for (int c = 0; c < g.n; ++c) {
      const T *x = xb + pairs[g.list_off + c].x * DD;
      for (int a = threadIdx.x; a < D; a += 64) {
        const double2 u = ld2(x + a * D + a);
        re += u.x;
        im += u.y;
      }
    }
  }
  re = wave_sum(re);
  im = wave_sum(im);
  if (threadIdx.x == 0) out[obase + ti * ostride + g.dst] = make_double2(re, im);
}

// ---------------------------------------------------------------------------
// Time-separable traces as small GEMMs over time on FP64 MFMA: for one
// (trace id, mask group) and one tile of nL x nR slice pairs,
//   out[a][b] = sum_(i,j) sum_{x,y} L_a,i[x,y] R_b,j[y,x],
// a 16 x 16 x (pairs * D^2) complex GEMM whose K index runs over the matrix
// elements.  The MFMA operands come straight from memory: for k-step s lane l
// supplies element e = 4s + l/16 of L_{a = l%16} and the partner element of
// R_{b = l%16} (index e itself when the right matrices are stored transposed,
// so four lanes read 64 contiguous bytes of each matrix).  Pairs that share a
// left matrix are merged first, sum_j tr(L_i R_j) = tr(L_i sum_j R_j): the
// right elements of a run of (up to two) equal i are added in registers and
// the run costs one pass of MFMAs.  Wave w takes the runs r with r % 4 == w;
// the 4 partial tiles are reduced through LDS.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mfma_c(d4 &cr, d4 &ci, double2 u, double2 v) {
  // (ux + i uy)(vx + i vy) = (ux vx - uy vy) + i (ux vy + uy vx)
  cr = __builtin_amdgcn_mfma_f64_16x16x4f64(u.x, v.x, cr, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-u.y, v.y, cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f64_16x16x4f64(u.x, v.y, ci, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f64_16x16x4f64(u.y, v.x, ci, 0, 0, 0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_trace_grid(const T *__restrict__ sX, const long long *__restrict__ Xq,
                                                    const T *__restrict__ sY, const long long *__restrict__ Yq,
                                                    const GroupItem *__restrict__ items, const int2 *__restrict__ pairs,
                                                    const GridTile *__restrict__ tiles, double2 *__restrict__ out,
                                                    long long obase, long long ostride, int D, int rt) {
  // Items in launch order, which the planner sorts by trace-graph component:
  // consecutive workgroups go to different XCDs, so all XCDs work on the same
  // few components at a time and their operand matrices stay in the shared
  // last-level cache (measured faster than giving each XCD its own range).
  const GroupItem g = items[blockIdx.x];
  const GridTile &Tl = tiles[blockIdx.y];
  __shared__ double2 part[4 * 256];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int DD = D * D, q = lane >> 4;
  const int nL = Tl.nL, nR = Tl.nR;
  const int la = min(lane & 15, nL - 1), lb = min(lane & 15, nR - 1);  // rows past the tile repeat its last
  const T *xrow = sX + (Tl.Lt[la] + Xq[g.qn]) * DD;
  const T *yrow = sY + (Tl.Rt[lb] + Yq[g.qn]) * DD;
  const int2 *pl = pairs + g.list_off;
  d4 cr = d4{0.0, 0.0, 0.0, 0.0}, ci = d4{0.0, 0.0, 0.0, 0.0};
  int run = 0;
  for (int c = 0; c < g.n; ++run) {
    const int i = pl[c].x;
    const bool two = c + 1 < g.n && pl[c + 1].x == i;
    if (run % 4 == wave) {
      const T *x = xrow + (long long)i * DD + q;
      const T *y0 = yrow + (long long)pl[c].y * DD;
      const T *y1 = yrow + (long long)pl[two ? c + 1 : c].y * DD;
      if (rt) {
        y0 += q;
        y1 += q;
        if (two) {
#pragma unroll 4
          for (int s = 0; s < DD / 4; ++s) mfma_c(cr, ci, ld2(x + 4 * s), cadd(ld2(y0 + 4 * s), ld2(y1 + 4 * s)));
        } else {
#pragma unroll 4
          for (int s = 0; s < DD / 4; ++s) mfma_c(cr, ci, ld2(x + 4 * s), ld2(y0 + 4 * s));
        }
      } else {
        for (int s = 0; s < DD / 4; ++s) {
          const int e = 4 * s + q, er = (e % D) * D + e / D;
          mfma_c(cr, ci, ld2(x + 4 * s), two ? cadd(ld2(y0 + er), ld2(y1 + er)) : ld2(y0 + er));
        }
      }
    }
    c += two ? 2 : 1;
  }
  // reduce the 4 waves' partial 16 x 16 tiles
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int a = (lane >> 4) + 4 * i, b = lane & 15;
    part[wave * 256 + a * 16 + b] = make_double2(cr[i], ci[i]);
  }
  __syncthreads();
  {
    const int a = tid / 16, b = tid % 16;
    double2 s = part[tid];
    s = cadd(s, part[256 + tid]);
    s = cadd(s, part[512 + tid]);
    s = cadd(s, part[768 + tid]);
    const int sp = (a < nL && b < nR) ? Tl.sidx[a][b] : -1;
    if (sp >= 0) out[obase + sp * ostride + g.dst] = s;
  }
}

// Right-stationary grid traces: one workgroup per (group of items that read
// the same right entry, slice-pair tile), groups in plan order per XCD.  For each run the right element
// (or the sum of the two right elements of a partner pair) is loaded once and
// multiplied into the tile of every item of the group that has the run, with
// that item's left element: per k-step one or two right loads and one left
// load per item feed four MFMAs per item.  Wave w takes the runs r with
// r % 4 == w; the 4 partial tiles of each item are reduced through LDS.
// Three waves per SIMD (164 registers): measured 113 ms per 32^3 x 64
// configuration, 142 ms when the compiler is left to use more registers for
// two waves per SIMD.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_trace_grid_r(const T *__restrict__ sX, const T *__restrict__ sY,
                                                      const GridRGroup *__restrict__ groups,
                                                      const int *__restrict__ runs, const GridTile *__restrict__ tiles,
                                                      double2 *__restrict__ out, long long obase, long long ostride,
                                                      int D, int rt) {
  // Workgroups are dealt round-robin to the 8 XCDs, each with its own L2:
  // XCD x runs the x-th contiguous eighth of the groups, in which
  // neighbouring groups read the same left entries.
  const int NG = gridDim.x, per = NG / 8, rem = NG % 8, x = blockIdx.x % 8, jx = blockIdx.x / 8;
  const int gi = x < rem ? x * (per + 1) + jx : rem * (per + 1) + (x - rem) * per + jx;
  const GridRGroup &G = groups[gi];
  const GridTile &Tl = tiles[blockIdx.y];
  __shared__ double2 part[4 * 256];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int DD = D * D, q = lane >> 4;
  const int nL = Tl.nL, nR = Tl.nR, nitem = G.nitem, nrun = G.nrun;
  const int la = min(lane & 15, nL - 1), lb = min(lane & 15, nR - 1);  // rows past the tile repeat its last
  const T *xrow = sX + Tl.Lt[la] * DD;
  const T *yrow = sY + (Tl.Rt[lb] + G.yq) * DD;
  const int *rr0 = runs + (long long)G.run_off * kGridRun;
  d4 cr[kGridIPW], ci[kGridIPW];
#pragma unroll
  for (int w = 0; w < kGridIPW; ++w) cr[w] = ci[w] = d4{0.0, 0.0, 0.0, 0.0};
  for (int r = wave; r < nrun; r += 4) {
    const int *rr = rr0 + r * kGridRun;
    const int j1 = rr[0], j2 = rr[1];
    int xo[kGridIPW];
#pragma unroll
    for (int w = 0; w < kGridIPW; ++w) xo[w] = rr[2 + w];
    const T *y0 = yrow + (long long)j1 * DD, *y1 = yrow + (long long)(j2 >= 0 ? j2 : j1) * DD;
    const T *xw[kGridIPW];
#pragma unroll
    for (int w = 0; w < kGridIPW; ++w) xw[w] = xrow + (long long)max(xo[w], 0) * DD + q;
    if (rt) {
      y0 += q;
      y1 += q;
#pragma unroll 2
      for (int s = 0; s < DD / 4; ++s) {
        const double2 v = j2 >= 0 ? cadd(ld2(y0 + 4 * s), ld2(y1 + 4 * s)) : ld2(y0 + 4 * s);
#pragma unroll
        for (int w = 0; w < kGridIPW; ++w)
          if (xo[w] >= 0) mfma_c(cr[w], ci[w], ld2(xw[w] + 4 * s), v);
      }
    } else {
      for (int s = 0; s < DD / 4; ++s) {
        const int e = 4 * s + q, er = (e % D) * D + e / D;
        const double2 v = j2 >= 0 ? cadd(ld2(y0 + er), ld2(y1 + er)) : ld2(y0 + er);
#pragma unroll
        for (int w = 0; w < kGridIPW; ++w)
          if (xo[w] >= 0) mfma_c(cr[w], ci[w], ld2(xw[w] + 4 * s), v);
      }
    }
  }
  // reduce the 4 waves' partial 16 x 16 tiles, item by item
  const int a = tid / 16, b = tid % 16;
  const int sp = (a < nL && b < nR) ? Tl.sidx[a][b] : -1;
#pragma unroll
  for (int w = 0; w < kGridIPW; ++w) {
    if (w >= nitem) break;
    if (w > 0) __syncthreads();  // the previous item's tile has been read
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ra = (lane >> 4) + 4 * i, rb = lane & 15;
      part[wave * 256 + ra * 16 + rb] = make_double2(cr[w][i], ci[w][i]);
    }
    __syncthreads();
    double2 v = part[tid];
    v = cadd(v, part[256 + tid]);
    v = cadd(v, part[512 + tid]);
    v = cadd(v, part[768 + tid]);
    if (sp >= 0) out[obase + sp * ostride + G.dst[w]] = v;
  }
}

// GEMM form of the slice-pair tile traces: one workgroup per tile of
// kGemmT x kGemmT (left entry, right entry) pairs, i.e. a 64 x 64 output
// block, rows (entry r, left time a), columns (entry c, right time b).  K runs
// over (run, matrix element) in chunks of 16 elements of one run: each thread
// loads 4 elements of one row of L and of one column of R (the two right
// matrices of the run added), the chunk goes through LDS, and the four waves
// (2 x 2, 32 x 32 each) multiply it with FP64 MFMA by three real products
// (Gauss).  The next chunk's loads are issued before the current chunk's
// MFMAs.  Every left element feeds 64 columns and every right element 64 rows.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_trace_gemm(
    const T *__restrict__ sX, const T *__restrict__ sY, const GridGTile *__restrict__ tiles,
    const int *__restrict__ runs, int nrun, const GridTile *__restrict__ times, double2 *__restrict__ out,
    long long obase, long long ostride, int D, int rt, double2 *__restrict__ part) {
  constexpr int KC = 16;
  // K split over gridDim.y parts (runs r0 .. r1); with more than one part
  // the partial tiles go to part[] and k_gemm_reduce sums them in order.
  // XCD x (workgroups are dealt to the XCDs round-robin in dispatch order)
  // runs the x-th contiguous eighth of the (part, tile) sequence, so that the
  // tiles of one component, which share left rows and right columns, meet in
  // the same L2 at the same K part.
  const int NB = gridDim.x * gridDim.y, lin = blockIdx.y * gridDim.x + blockIdx.x;
  const int per = NB / 8, rem = NB % 8, xcd = lin % 8, jx = lin / 8;
  const int lb = xcd < rem ? xcd * (per + 1) + jx : rem * (per + 1) + (xcd - rem) * per + jx;
  const int tile = lb % gridDim.x, kp = lb / gridDim.x, KP = gridDim.y;
  const GridGTile &G = tiles[tile];
  const int r0 = int((long long)nrun * kp / KP), r1 = int((long long)nrun * (kp + 1) / KP);
  const GridTile &Tm = times[0];
  __shared__ double2 As[64][KC + 1];
  __shared__ double2 Bs[64][KC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  const int DD = D * D, nL = Tm.nL, nR = Tm.nR;
  // this thread's row of L / column of R and its 4 elements of a chunk
  const int lr = tid >> 2, lk = (tid & 3) * 4;
  const int er = min(lr >> 4, G.nE - 1), fc = min(lr >> 4, G.nF - 1);
  const int ta = min(lr & 15, nL - 1), tb = min(lr & 15, nR - 1);  // rows past the time tile repeat its last
  const T *xrow = sX + (Tm.Lt[ta] + G.xq[er]) * DD;
  const T *yrow = sY + (Tm.Rt[tb] + G.yq[fc]) * DD;
  const int cpr = DD / KC;  // chunks per run
  const int nchunk = (r1 - r0) * cpr;
  double2 av[4], bv[4];
  auto load = [&](int c) {
    const int sidx = r0 + c / cpr, e0 = (c % cpr) * KC + lk;
    const int i = runs[3 * sidx], j1 = runs[3 * sidx + 1], j2 = runs[3 * sidx + 2];
    const T *xa = xrow + (long long)i * DD;
    const T *y1 = yrow + (long long)j1 * DD, *y2 = yrow + (long long)(j2 >= 0 ? j2 : j1) * DD;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = e0 + u, er2 = rt ? e : (e % D) * D + e / D;
      av[u] = ld2(xa + e);
      bv[u] = j2 >= 0 ? cadd(ld2(y1 + er2), ld2(y2 + er2)) : ld2(y1 + er2);
    }
  };
  // (ax + i ay)(bx + i by): P1 = ax bx, P2 = ay by, P3 = (ax + ay)(bx + by);
  // re = P1 - P2, im = P3 - P1 - P2
  d4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) p1[a][b] = p2[a][b] = p3[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  if (nchunk > 0) load(0);
  for (int c = 0; c < nchunk; ++c) {
    __syncthreads();  // the previous chunk has been read
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      As[lr][lk + u] = av[u];
      Bs[lr][lk + u] = bv[u];
    }
    __syncthreads();
    if (c + 1 < nchunk) load(c + 1);
#pragma unroll
    for (int kq = 0; kq < KC; kq += 4) {
      const int kr = kq + (lane >> 4);
      const double2 a2[2] = {As[wm * 32 + (lane & 15)][kr], As[wm * 32 + 16 + (lane & 15)][kr]};
      const double2 b2[2] = {Bs[wn * 32 + (lane & 15)][kr], Bs[wn * 32 + 16 + (lane & 15)][kr]};
      const double as[2] = {a2[0].x + a2[0].y, a2[1].x + a2[1].y};
      const double bs[2] = {b2[0].x + b2[0].y, b2[1].x + b2[1].y};
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
          p1[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[mi].x, b2[ni].x, p1[mi][ni], 0, 0, 0);
          p2[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[mi].y, b2[ni].y, p2[mi][ni], 0, 0, 0);
          p3[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[mi], bs[ni], p3[mi][ni], 0, 0, 0);
        }
    }
  }
  // rows (lane >> 4) + 4 i of each 16 x 16 block, column lane & 15
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = wm * 32 + mi * 16 + (lane >> 4) + 4 * i, col = wn * 32 + ni * 16 + (lane & 15);
        if (KP > 1) {
          const double q1 = p1[mi][ni][i], q2 = p2[mi][ni][i];
          part[((long long)tile * KP + kp) * 4096 + row * 64 + col] =
              make_double2(q1 - q2, p3[mi][ni][i] - q1 - q2);
          continue;
        }
        const int r = row >> 4, a = row & 15, q = col >> 4, b = col & 15;
        if (r >= G.nE || q >= G.nF || a >= nL || b >= nR) continue;
        const long long dst = G.dst[r][q];
        if (dst < 0) continue;
        const int sp = Tm.sidx[a][b];
        if (sp < 0) continue;
        const double q1 = p1[mi][ni][i], q2 = p2[mi][ni][i];
        out[obase + sp * ostride + dst] = make_double2(q1 - q2, p3[mi][ni][i] - q1 - q2);
      }
}

// Sum of the K parts of each GEMM tile element in part order, stored into its
// item like the one-part kernel does.
__global__ void k_gemm_reduce(const double2 *__restrict__ part, int KP, const GridGTile *__restrict__ tiles,
                              int ntiles, const GridTile *__restrict__ times, double2 *__restrict__ out, long long obase,
                              long long ostride) {
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (id >= (long long)ntiles * 4096) return;
  const int t = int(id >> 12), el = int(id & 4095), row = el >> 6, col = el & 63;
  const GridGTile &G = tiles[t];
  const GridTile &Tm = times[0];
  const int r = row >> 4, a = row & 15, q = col >> 4, b = col & 15;
  if (r >= G.nE || q >= G.nF || a >= Tm.nL || b >= Tm.nR) return;
  const long long dst = G.dst[r][q];
  const int sp = Tm.sidx[a][b];
  if (dst < 0 || sp < 0) return;
  double2 v = part[(long long)t * KP * 4096 + el];
  for (int k = 1; k < KP; ++k) v = cadd(v, part[((long long)t * KP + k) * 4096 + el]);
  out[obase + sp * ostride + dst] = v;
}

// ---------------------------------------------------------------------------
// resolve_request{1,2,3} (reference: Diagram.cpp:9-94, DilutedTrace.cpp:5-82)
// on mask-group sums: mean, sum_g L_g sum_{h: m_g & m_h = 0} R_h, and the
// 3-way form that rejects only masks shared by all three traces
// (DilutedTrace.cpp:63).  One thread per (correlator, up to 8 slice pairs).  When the
// traces carry subset sums (zbase >= 0), the 2- and 3-way forms are evaluated
// by Moebius inversion: sum_S (-1)^|S| Z0(S) Z1(S) (Z2(S)) over the subset
// list, the same sum re-associated (nz terms instead of ngroup^ntr).
// ---------------------------------------------------------------------------
constexpr int kZetaRegs = 16;  // subset lists up to this length keep their sums in registers
__global__ void k_zeta(const double2 *__restrict__ gv, long long gbase, long long gstride,
                       const uint16_t *__restrict__ gm, const ZItem *__restrict__ items, int nitems,
                       const uint16_t *__restrict__ zmask, int nz, double2 *__restrict__ zv, long long zbase,
                       long long zstride, int nsp) {
  // one thread per (slice pair, item), every subset in turn: subset-major
  // output, so consecutive threads write consecutive items of one subset
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (id >= (long long)nsp * nitems) return;
  const int it = int(id % nitems), s = int(id / nitems);
  const ZItem I = items[it];
  const double2 *v = gv + gbase + s * gstride + I.goff;
  const uint16_t *m = gm + I.gmask_off;
  double2 *out = zv + zbase + s * zstride + it;
  if (nz <= kZetaRegs) {
    // every group sum read once, the subset sums in registers (each in the
    // same order of groups as below)
    double2 z[kZetaRegs];
#pragma unroll
    for (int q = 0; q < kZetaRegs; ++q) z[q] = make_double2(0.0, 0.0);
    for (int g = 0; g < I.ngroup; ++g) {
      const double2 x = v[g];
      const uint16_t mg = m[g];
#pragma unroll
      for (int q = 0; q < kZetaRegs; ++q)
        if (q < nz && (mg & zmask[q]) == zmask[q]) z[q] = cadd(z[q], x);
    }
#pragma unroll
    for (int q = 0; q < kZetaRegs; ++q)
      if (q < nz) out[(long long)q * nitems] = z[q];
    return;
  }
  for (int q = 0; q < nz; ++q) {
    const uint16_t S = zmask[q];
    double2 z = make_double2(0.0, 0.0);
    for (int g = 0; g < I.ngroup; ++g)
      if ((m[g] & S) == S) z = cadd(z, v[g]);
    out[(long long)q * nitems] = z;
  }
}

// Neumaier-Kahan accumulator of KahanAccumulator<Complex>::operator+=
// (include/KahanAccumulator.hpp:23-32): the branch compares complex
// magnitudes, value() = sum + compensation.
struct KahanC {
  double2 s, c;
  __device__ __forceinline__ KahanC() : s(make_double2(0.0, 0.0)), c(make_double2(0.0, 0.0)) {}
  __device__ __forceinline__ void add(double2 x) {
    const double2 t = cadd(s, x);
    if (hypot(s.x, s.y) >= hypot(x.x, x.y)) {
      c.x += (s.x - t.x) + x.x;
      c.y += (s.y - t.y) + x.y;
    } else {
      c.x += (x.x - t.x) + s.x;
      c.y += (x.y - t.y) + s.y;
    }
    s = t;
  }
  __device__ __forceinline__ double2 value() const { return cadd(s, c); }
};

// One (slice pair, correlator): resolve_request{1,2,3} (Diagram.cpp:9-94).
// Plain sums in the default form; with kahan != 0 every group is one trace
// value (the planner's per-element form) and the sums run as in the
// reference's SLAPH_KAHAN build: the mean of resolve_request1 through
// std::accumulate into a Kahan accumulator, and right_sum and result of
// inner_product (DilutedTrace.cpp:5-82) as Kahan accumulators, in the
// reference's element order.  The result is divided by the number of
// admitted combinations, as the reference does.
__global__ void k_resolve(const double2 *__restrict__ gv, const uint16_t *__restrict__ gm,
                          const CorrJob *__restrict__ jobs, int ncorr, int nsp, const double2 *__restrict__ zv,
                          const uint16_t *__restrict__ zmask, int nz, double2 *__restrict__ cval, int kahan, int spt) {
  // XCD x (workgroups are dealt to the XCDs round-robin) runs the x-th
  // contiguous eighth of the (slice pair, correlator) sequence, so each XCD's
  // L2 holds the trace values of its own slice pairs only.
  const int NB = gridDim.x, per = NB / 8, rem = NB % 8, xcd = blockIdx.x % 8, jx = blockIdx.x / 8;
  const int lb = xcd < rem ? xcd * (per + 1) + jx : rem * (per + 1) + (xcd - rem) * per + jx;
  // one thread per (correlator, spt slice pairs): the job record is read
  // once for all of them, and for each slice pair the lanes of a wave write
  // consecutive correlators
  const long long id = lb * (long long)blockDim.x + threadIdx.x;
  const int nsb = (nsp + spt - 1) / spt;
  if (id >= (long long)ncorr * nsb) return;
  const int corr = int(id % ncorr), sb = int(id / ncorr);
  const CorrJob J = jobs[corr];
  for (int s = sb * spt; s < min(nsp, (sb + 1) * spt); ++s) {
  const double2 *v0 = gv + J.tr[0].base + s * J.tr[0].stride;
  double2 res = make_double2(0.0, 0.0);
  if (kahan) {
    KahanC acc;
    if (J.ntr == 1) {
      for (int g = 0; g < J.tr[0].ngroup; ++g) acc.add(v0[g]);
    } else if (J.ntr == 2) {
      const double2 *v1 = gv + J.tr[1].base + s * J.tr[1].stride;
      const uint16_t *m0 = gm + J.tr[0].gmask_off, *m1 = gm + J.tr[1].gmask_off;
      for (int g = 0; g < J.tr[0].ngroup; ++g) {
        KahanC rs;
        for (int h = 0; h < J.tr[1].ngroup; ++h)
          if (!(m0[g] & m1[h])) rs.add(v1[h]);
        acc.add(cmul(v0[g], rs.value()));
      }
    } else {
      const double2 *v1 = gv + J.tr[1].base + s * J.tr[1].stride;
      const double2 *v2 = gv + J.tr[2].base + s * J.tr[2].stride;
      const uint16_t *m0 = gm + J.tr[0].gmask_off, *m1 = gm + J.tr[1].gmask_off, *m2 = gm + J.tr[2].gmask_off;
      for (int g = 0; g < J.tr[0].ngroup; ++g)
        for (int h = 0; h < J.tr[1].ngroup; ++h) {
          const uint16_t mgh = m0[g] & m1[h];
          KahanC rs;
          for (int k = 0; k < J.tr[2].ngroup; ++k)
            if (!(mgh & m2[k])) rs.add(v2[k]);
          acc.add(cmul(cmul(v0[g], v1[h]), rs.value()));
        }
    }
    res = acc.value();
  } else if (J.ntr == 1) {
    for (int g = 0; g < J.tr[0].ngroup; ++g) res = cadd(res, v0[g]);
  } else if (J.tr[0].zbase >= 0) {
    // subset-major sums: the lanes of a wave (jobs ordered by their traces)
    // mostly share their first two traces, read then through scalar loads,
    // and read neighbouring items of the last
    const long long b0 = J.tr[0].zbase + s * J.tr[0].zstride, b1 = J.tr[1].zbase + s * J.tr[1].zstride;
    const double2 *z2 = J.ntr == 3 ? zv + J.tr[2].zbase + s * J.tr[2].zstride : nullptr;
    const int q0 = J.tr[0].zq, q1 = J.tr[1].zq, q2 = J.ntr == 3 ? J.tr[2].zq : 0;
    auto rfl = [](long long v) {
      return (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(int(v >> 32)) << 32) |
                         (unsigned)__builtin_amdgcn_readfirstlane(int(v)));
    };
    const long long u0 = rfl(b0), u1 = rfl(b1);
    const int uq0 = __builtin_amdgcn_readfirstlane(q0), uq1 = __builtin_amdgcn_readfirstlane(q1);
    auto term = [&](double2 t, int q) {
      if (z2) t = cmul(t, z2[q * q2]);
      if (__popc(zmask[q]) & 1) {
        res.x -= t.x;
        res.y -= t.y;
      } else {
        res = cadd(res, t);
      }
    };
    if (__all(b0 == u0 && b1 == u1 && q0 == uq0 && q1 == uq1)) {
      const double2 *z0 = zv + u0, *z1 = zv + u1;
#pragma unroll 4
      for (int q = 0; q < nz; ++q) term(cmul(z0[q * uq0], z1[q * uq1]), q);
    } else {
      const double2 *z0 = zv + b0, *z1 = zv + b1;
#pragma unroll 4
      for (int q = 0; q < nz; ++q) term(cmul(z0[q * q0], z1[q * q1]), q);
    }
  } else if (J.ntr == 2) {
    const double2 *v1 = gv + J.tr[1].base + s * J.tr[1].stride;
    const uint16_t *m0 = gm + J.tr[0].gmask_off, *m1 = gm + J.tr[1].gmask_off;
    for (int g = 0; g < J.tr[0].ngroup; ++g) {
      double2 rs = make_double2(0.0, 0.0);
      for (int h = 0; h < J.tr[1].ngroup; ++h)
        if (!(m0[g] & m1[h])) rs = cadd(rs, v1[h]);
      res = cadd(res, cmul(v0[g], rs));
    }
  } else {
    const double2 *v1 = gv + J.tr[1].base + s * J.tr[1].stride;
    const double2 *v2 = gv + J.tr[2].base + s * J.tr[2].stride;
    const uint16_t *m0 = gm + J.tr[0].gmask_off, *m1 = gm + J.tr[1].gmask_off, *m2 = gm + J.tr[2].gmask_off;
    for (int g = 0; g < J.tr[0].ngroup; ++g)
      for (int h = 0; h < J.tr[1].ngroup; ++h) {
        const uint16_t mgh = m0[g] & m1[h];
        double2 rs = make_double2(0.0, 0.0);
        for (int k = 0; k < J.tr[2].ngroup; ++k)
          if (!(mgh & m2[k])) rs = cadd(rs, v2[k]);
        res = cadd(res, cmul(cmul(v0[g], v1[h]), rs));
      }
  }
  cval[(long long)s * ncorr + J.out] = make_double2(res.x / J.n, res.y / J.n);
  }
}

// The Moebius form alone (every job of the launch has subset sums and no
// Kahan form): the same sums as k_resolve's with a fraction of its registers,
// so that more waves hide the load latency.
__global__ void k_resolve_z(const CorrJob *__restrict__ jobs, int ncorr, int nsp, const double2 *__restrict__ zv,
                            const uint16_t *__restrict__ zmask, int nz, double2 *__restrict__ cval, int spt) {
  const int NB = gridDim.x, per = NB / 8, rem = NB % 8, xcd = blockIdx.x % 8, jx = blockIdx.x / 8;
  const int lb = xcd < rem ? xcd * (per + 1) + jx : rem * (per + 1) + (xcd - rem) * per + jx;
  const long long id = lb * (long long)blockDim.x + threadIdx.x;
  const int nsb = (nsp + spt - 1) / spt;
  if (id >= (long long)ncorr * nsb) return;
  const int corr = int(id % ncorr), sb = int(id / ncorr);
  const CorrJob J = jobs[corr];
  auto rfl = [](long long v) {
    return (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(int(v >> 32)) << 32) |
                       (unsigned)__builtin_amdgcn_readfirstlane(int(v)));
  };
  const int q0 = J.tr[0].zq, q1 = J.tr[1].zq, q2 = J.ntr == 3 ? J.tr[2].zq : 0;
  const int uq0 = __builtin_amdgcn_readfirstlane(q0), uq1 = __builtin_amdgcn_readfirstlane(q1);
  for (int s = sb * spt; s < min(nsp, (sb + 1) * spt); ++s) {
    // subset-major sums: the lanes of a wave (jobs ordered by their traces)
    // mostly share their first trace, read then through scalar loads (the
    // second too when the whole wave shares it), and read neighbouring
    // items of the last
    const long long b0 = J.tr[0].zbase + s * J.tr[0].zstride, b1 = J.tr[1].zbase + s * J.tr[1].zstride;
    const double2 *z2 = J.ntr == 3 ? zv + J.tr[2].zbase + s * J.tr[2].zstride : nullptr;
    const long long u0 = rfl(b0), u1 = rfl(b1);
    double2 res = make_double2(0.0, 0.0);
    auto sum = [&](const double2 *z0, int s0, const double2 *z1, int s1) {
#pragma unroll 4
      for (int q = 0; q < nz; ++q) {
        double2 t = cmul(z0[q * s0], z1[q * s1]);
        if (z2) t = cmul(t, z2[q * q2]);
        if (__popc(zmask[q]) & 1) {
          res.x -= t.x;
          res.y -= t.y;
        } else {
          res = cadd(res, t);
        }
      }
    };
    if (__all(b0 == u0 && q0 == uq0)) {
      if (__all(b1 == u1 && q1 == uq1))
        sum(zv + u0, uq0, zv + u1, uq1);
      else
        sum(zv + u0, uq0, zv + b1, q1);
    } else {
      sum(zv + b0, q0, zv + b1, q1);
    }
    cval[(long long)s * ncorr + J.out] = make_double2(res.x / J.n, res.y / J.n);
  }
}

// correlator_[t][i] += c (reference: Diagram::assemble, Diagram.cpp:96-114).
// The reference gives every slice pair a fresh accumulator holding its one
// value and merges it with KahanAccumulator::operator+=(KahanAccumulator)
// (KahanAccumulator.hpp:40-44), which adds sums and compensations separately:
// the compensation stays zero, so across slice pairs the sum is plain in both
// builds.  One thread per (time separation, correlator): the slice pairs of
// that separation are added in ascending order into a register copy of the
// sum, the same sequence of additions as one pass over all slice pairs.
__global__ void k_accumulate(const double2 *__restrict__ cval, const int *__restrict__ sidx,
                             const int4 *__restrict__ tgroups, int ngroups, int ncorr, double2 *__restrict__ sum) {
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (id >= (long long)ngroups * ncorr) return;
  const int corr = int(id % ncorr);
  const int4 G = tgroups[id / ncorr];
  const long long at = (long long)G.x * ncorr + corr;
  double2 acc = sum[at];
  for (int k = 0; k < G.z; ++k) acc = cadd(acc, cval[(long long)sidx[G.y + k] * ncorr + corr]);
  sum[at] = acc;
}

__global__ void k_finalize(const double2 *__restrict__ sum, int ncorr, int Lt, const int *__restrict__ orig,
                           double2 *__restrict__ out) {
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (id >= (long long)ncorr * Lt) return;
  const int t = int(id % Lt), i = int(id / Lt);
  const long long at = (long long)t * ncorr + i;
  out[(long long)(orig ? orig[i] : i) * Lt + t] = make_double2(sum[at].x / double(Lt), sum[at].y / double(Lt));
}

template <typename T>
__global__ void k_to_storage(const double2 *__restrict__ in, T *__restrict__ out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    st2(out + i, in[i]);
}

__global__ void k_scale(double2 *x, long long n, double f) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    x[i].x *= f;
    x[i].y *= f;
  }
}

// splitmix64 finaliser as a counter-based generator: value i of stream `seed`.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(unsigned long long z) { return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

__global__ void k_fill_gauss(double2 *x, long long n, unsigned long long seed) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const unsigned long long h = mix64(seed ^ mix64((unsigned long long)i));
    const double u1 = u01(h), u2 = u01(mix64(h));
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(6.283185307179586 * u2, &sn, &cs);
    x[i] = make_double2(r * cs * 0.7071067811865476, r * sn * 0.7071067811865476);
  }
}

__global__ void k_fill_z2z2(double2 *x, long long n, unsigned long long seed) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const unsigned long long h = mix64(seed ^ mix64((unsigned long long)i));
    const double a = (h & 1) ? 0.7071067811865476 : -0.7071067811865476;
    const double b = (h & 2) ? 0.7071067811865476 : -0.7071067811865476;
    x[i] = make_double2(a, b);
  }
}

// C = A R with A: K x N, R: N x N small, one thread per row segment.
__global__ __launch_bounds__(256) void k_gemm_nn_small(const double2 *__restrict__ A, const double2 *__restrict__ R,
                                                       double2 *__restrict__ C, long long K, int N) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (k >= K) return;
  double2 s = make_double2(0.0, 0.0);
  for (int j = 0; j < N; ++j) cfma(s, A[(long long)j * K + k], R[(long long)c * N + j]);
  C[(long long)c * K + k] = s;
}

bool g_const_ready[64] = {false};

void upload_gamma_constants() {
  int dev = 0;
  SLAPH_HIP(hipGetDevice(&dev));
  if (dev < 64 && g_const_ready[dev]) return;
  int rows[16][4];
  double2 vals[16][4];
  // Columns k of gamma g: value at row rows[g][k] (tmLQCD basis products,
  // same derivation as the host's gamma_table()).
  double tmp[16][4][2];
  host_gamma_tables(rows, tmp);
  for (int g = 0; g < 16; ++g)
    for (int k = 0; k < 4; ++k) vals[g][k] = make_double2(tmp[g][k][0], tmp[g][k][1]);
  SLAPH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gamma_row), rows, sizeof rows));
  SLAPH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gamma_val), vals, sizeof vals));
  if (dev < 64) g_const_ready[dev] = true;
}

dim3 grid1d(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g > 65535LL * 16) g = 65535LL * 16;
  if (g < 1) g = 1;
  return dim3(unsigned(g));
}

}  // namespace

// Calls f with a value of the element type of storage form `form`
// (0 = double2, 1 = float2, 2 = d48c).
template <typename F>
void with_storage(int form, F &&f) {
  switch (form) {
    case 0:
      f(double2{});
      break;
    case 1:
      f(float2{});
      break;
    case 2:
      f(d48c{});
      break;
    default:
      throw std::runtime_error("unknown storage form " + std::to_string(form));
  }
}

// Job table of the allocation-free A^H B entry (slc_ahb_device_noalloc): one
// fixed array per GPU instead of a per-call allocation.  In host code the
// name `scratch` is only a stand-in for the symbol; the device address comes
// from hipGetSymbolAddress (ahb_noalloc_table) and is the only one passed on.
static __device__ __align__(256) AhBJob scratch[kAhbNoallocJobs];

AhBJob *ahb_noalloc_table(int device) {
  static std::mutex mu;
  static AhBJob *addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) throw std::out_of_range("device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    int cur = -1;
    SLAPH_HIP(hipGetDevice(&cur));
    if (cur != device) SLAPH_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) SLAPH_HIP(hipSetDevice(cur));
    SLAPH_HIP(e);
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the A^H B job table");
    addr[device] = static_cast<AhBJob *>(p);
  }
  return addr[device];
}

void launch_ahb(const AhBJob *d_jobs, int njobs, int M, int N, int K, hipStream_t s) {
  if (njobs <= 0) return;
  if (njobs > 65535) throw std::runtime_error("launch_ahb: too many jobs in one launch");
  dim3 grid((M + AHB_TM - 1) / AHB_TM, (N + AHB_TN - 1) / AHB_TN, njobs);
  hipLaunchKernelGGL(k_ahb, grid, dim3(256), 0, s, d_jobs, M, N, K);
  SLAPH_HIP(hipGetLastError());
}

void launch_ahb_grouped(const AhBGroup *d_groups, int ngroups, int max_jobs, const AhBRow *d_rows, int rpj, int N,
                        int K, long long lda, long long ldb, long long ldc, hipStream_t s) {
  if (ngroups <= 0 || max_jobs <= 0) return;
  if (ngroups > 65535) throw std::runtime_error("launch_ahb_grouped: too many groups in one launch");
  dim3 grid((max_jobs * rpj + AHB_TM - 1) / AHB_TM, (N + AHB_TN - 1) / AHB_TN, ngroups);
  hipLaunchKernelGGL(k_ahb_grouped, grid, dim3(256), 0, s, d_groups, d_rows, rpj, N, K, lda, ldb, ldc);
  SLAPH_HIP(hipGetLastError());
}

void launch_vdv(const VdvJob *d_jobs, int njobs, int N, long long K, int f32, hipStream_t s) {
  if (njobs <= 0) return;
  if (njobs > 65535) throw std::runtime_error("launch_vdv: too many jobs in one launch");
  const int tiles = (N + VDV_T - 1) / VDV_T;
  if (f32)
    hipLaunchKernelGGL(k_vdv<true>, dim3(tiles * tiles, njobs), dim3(256), 0, s, d_jobs, N, K, tiles);
  else
    hipLaunchKernelGGL(k_vdv<false>, dim3(tiles * tiles, njobs), dim3(256), 0, s, d_jobs, N, K, tiles);
  SLAPH_HIP(hipGetLastError());
}

void launch_vdv_split(const VdvJob *d_jobs, int njobs, int N, long long K, int form, hipStream_t s) {
  if (njobs <= 0) return;
  if (njobs > 65535) throw std::runtime_error("launch_vdv_split: too many jobs in one launch");
  if (form != 2 && form != 3) throw std::runtime_error("launch_vdv_split: form must be 2 (split BF16) or 3 (Ozaki)");
  const int tiles = (N + VDV_T - 1) / VDV_T;
  if (form == 2)
    hipLaunchKernelGGL(k_vdv_split<0>, dim3(tiles * tiles, njobs), dim3(256), 0, s, d_jobs, N, K, tiles);
  else
    hipLaunchKernelGGL(k_vdv_split<1>, dim3(tiles * tiles, njobs), dim3(256), 0, s, d_jobs, N, K, tiles);
  SLAPH_HIP(hipGetLastError());
}

long long ahb_big_part_elems(int njobs, int N, long long K) {
  const int tiles = (N + kBigT - 1) / kBigT;
  return (long long)njobs * ahb_big_parts(njobs, N, K) * tiles * tiles * kBigT * kBigT;
}

int ahb_big_parts(int njobs, int N, long long K) {
  // about 8 workgroups per compute unit of 256, each part at least 64 chunks
  const int tiles = (N + kBigT - 1) / kBigT;
  const long long want = (8LL * 256 + (long long)tiles * tiles * njobs - 1) / ((long long)tiles * tiles * njobs);
  const long long most = std::max(1LL, K / (64LL * kBigKC));
  return int(std::max(1LL, std::min({want, most, 64LL})));
}

void launch_ahb_big(const AhbBigJob *d_jobs, int njobs, int N, long long K, double2 *part, hipStream_t s) {
  if (njobs <= 0 || N <= 0 || K <= 0) return;
  if (njobs > 65535) throw std::runtime_error("launch_ahb_big: too many jobs");
  const int tiles = (N + kBigT - 1) / kBigT, parts = ahb_big_parts(njobs, N, K);
  const long long chunks = (K + kBigKC - 1) / kBigKC;
  const long long kper = ((chunks + parts - 1) / parts) * kBigKC;
  hipLaunchKernelGGL(k_ahb_big, dim3(unsigned(tiles * tiles), unsigned(parts), unsigned(njobs)), dim3(256), 0, s,
                     d_jobs, N, K, tiles, kper, part);
  SLAPH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_ahb_big_reduce, grid1d((long long)njobs * N * N, 256), dim3(256), 0, s, d_jobs, njobs, N,
                     tiles, parts, part);
  SLAPH_HIP(hipGetLastError());
}

void launch_colexp(const double2 *V, long long K, int N, int *e, hipStream_t s) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_colexp, dim3(N), dim3(256), 0, s, V, K, e);
  SLAPH_HIP(hipGetLastError());
}

void launch_vdv_ft(const double2 *Vr, long long vstride, FtLayout lay, int Lx, int Ly, int Lz, const int2 *tiles,
                   int ntp, int nt, FtPass const &pass, const double2 *ey, const double2 *zcs, double2 *U,
                   hipStream_t s) {
  if (nt <= 0 || ntp <= 0 || pass.nc <= 0) return;
  if (pass.nc > kFtMaxC || pass.nk < 0 || pass.nk > kFtMaxK)
    throw std::runtime_error("launch_vdv_ft: pass exceeds the combination limits");
  if (lay.ncol < 1) throw std::runtime_error("launch_vdv_ft: no columns");
  if (Lz > kFtMaxLz) throw std::runtime_error("launch_vdv_ft: Lz above kFtMaxLz");
  const long long blocks = (long long)nt * Lx * ntp;
  if (blocks > 0x7fffffffLL) throw std::runtime_error("launch_vdv_ft: grid too large");
  // Lz % 4 == 0: the incremental-address form; otherwise the general one
  const bool fast = (Lz & 3) == 0;
  if (pass.nc <= 9 && pass.nk <= 1) {
    if (fast)
      hipLaunchKernelGGL((k_vdv_ft<9, 1, true>), dim3(unsigned(blocks)), dim3(64), 0, s, Vr, vstride, lay, Lx, Ly,
                         Lz, tiles, ntp, pass, ey, zcs, U);
    else
      hipLaunchKernelGGL((k_vdv_ft<9, 1, false>), dim3(unsigned(blocks)), dim3(64), 0, s, Vr, vstride, lay, Lx, Ly,
                         Lz, tiles, ntp, pass, ey, zcs, U);
  } else {
    if (fast)
      hipLaunchKernelGGL((k_vdv_ft<kFtMaxC, kFtMaxK, true>), dim3(unsigned(blocks)), dim3(64), 0, s, Vr, vstride,
                         lay, Lx, Ly, Lz, tiles, ntp, pass, ey, zcs, U);
    else
      hipLaunchKernelGGL((k_vdv_ft<kFtMaxC, kFtMaxK, false>), dim3(unsigned(blocks)), dim3(64), 0, s, Vr, vstride,
                         lay, Lx, Ly, Lz, tiles, ntp, pass, ey, zcs, U);
  }
  SLAPH_HIP(hipGetLastError());
}

void launch_vdv_xsum(const double2 *U, int nc, int ntp, int nt, int Lx, const int2 *tiles, const FtTarget *targets,
                     int ntarget, const double2 *ex, int N, int t0, long long tstride, hipStream_t s) {
  if (nt <= 0 || ntp <= 0 || ntarget <= 0) return;
  if (nt > 65535) throw std::runtime_error("launch_vdv_xsum: too many timeslices in one launch");
  hipLaunchKernelGGL(k_vdv_xsum, dim3(ntp, nt), dim3(256), 0, s, U, nc, ntp, Lx, tiles, targets, ntarget, ex, N, t0,
                     tstride);
  SLAPH_HIP(hipGetLastError());
}

void launch_vrow(const double2 *V, double2 *Vr, long long K, int N, int Npad, int nt, hipStream_t s) {
  if (nt <= 0) return;
  if (Npad % 16 != 0 || Npad < N) throw std::runtime_error("launch_vrow: bad Npad");
  const long long kb = (K + 63) / 64;
  if (kb > 0x7fffffffLL || nt > 65535) throw std::runtime_error("launch_vrow: grid too large");
  hipLaunchKernelGGL(k_vrow, dim3(unsigned(kb), Npad / 16, nt), dim3(256), 0, s, V, Vr, K, N, Npad);
  SLAPH_HIP(hipGetLastError());
}

void launch_q0(const Q0Group *d_groups, int ngroups, const Q0Out *d_outs, int N, int dilE, hipStream_t s) {
  if (ngroups <= 0) return;
  upload_gamma_constants();
  size_t lds = sizeof(double2) * N * 4 * dilE;
  if (lds > 160 * 1024) throw std::runtime_error("Q0: N*4*dilE too large for LDS");
  hipLaunchKernelGGL(k_q0, dim3(ngroups), dim3(256), lds, s, d_groups, d_outs, N, dilE);
  SLAPH_HIP(hipGetLastError());
}

void launch_q1(const Q1Job *d_jobs, int njobs, int N, int dilE, long long prow, hipStream_t s) {
  if (njobs <= 0) return;
  upload_gamma_constants();
  size_t lds = sizeof(double2) * N * 4 * dilE;
  if (lds > 160 * 1024) throw std::runtime_error("Q1: N*4*dilE too large for LDS");
  hipLaunchKernelGGL(k_q1, dim3(njobs), dim3(256), lds, s, d_jobs, N, dilE, prow);
  SLAPH_HIP(hipGetLastError());
}

void launch_q2b(const Q2BJob *d_jobs, int njobs, int N, int dilE, long long prow, hipStream_t s) {
  if (njobs <= 0) return;
  upload_gamma_constants();
  if (dilE == 4 && N % 4 == 0)
    hipLaunchKernelGGL(k_q2b16, dim3((njobs + 3) / 4), dim3(256), 0, s, d_jobs, njobs, N, prow);
  else if (dilE == 6 && N % 4 == 0)
    hipLaunchKernelGGL(k_q2b_mfma<24>, dim3((njobs + 3) / 4), dim3(256), 0, s, d_jobs, njobs, N, prow);
  else if (dilE == 8 && N % 4 == 0)
    hipLaunchKernelGGL(k_q2b_mfma<32>, dim3((njobs + 3) / 4), dim3(256), 0, s, d_jobs, njobs, N, prow);
  else
    hipLaunchKernelGGL(k_q2b, dim3(njobs), dim3(256), 0, s, d_jobs, N, dilE, prow);
  SLAPH_HIP(hipGetLastError());
}

void launch_dagger(const double2 *in, double2 *out, int N, int count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_dagger, grid1d((long long)N * N * count, 256), dim3(256), 0, s, in, out, N, count);
  SLAPH_HIP(hipGetLastError());
}

void launch_zero(double2 *p, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero2, grid1d(n, 256), dim3(256), 0, s, p, n);
  SLAPH_HIP(hipGetLastError());
}

void launch_identity(double2 *out, int N, int count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_identity, grid1d((long long)N * N * count, 256), dim3(256), 0, s, out, N, count);
  SLAPH_HIP(hipGetLastError());
}

namespace {
template <typename T>
void launch_prod_t(const void *sX, const long long *Xt, const long long *Xq, const void *sY, const long long *Yt,
                   const long long *Yq, void *out, long long obase, long long ostride, const ComboProd *combos,
                   int ncombo, int ntime, int D, int transposed, int generic, hipStream_t s) {
  const T *x = static_cast<const T *>(sX), *y = static_cast<const T *>(sY);
  T *o = static_cast<T *>(out);
  if ((D == 24 || D == 32) && !generic && !std::getenv("SLAPH_PROD_GENERIC")) {
    const dim3 grid((ncombo + 3) / 4, ntime);
    if ((long long)grid.x * grid.y > 0x7fffffffLL) throw std::runtime_error("launch_prod: grid too large");
#define SLAPH_PROD_MFMA(DD)                                                                                       \
  if (transposed)                                                                                                 \
    hipLaunchKernelGGL((k_prod_mfma<T, DD, true>), grid, dim3(256), 0, s, x, Xt, Xq, y, Yt, Yq, o, obase, ostride, \
                       combos, ncombo);                                                                           \
  else                                                                                                            \
    hipLaunchKernelGGL((k_prod_mfma<T, DD, false>), grid, dim3(256), 0, s, x, Xt, Xq, y, Yt, Yq, o, obase, ostride, \
                       combos, ncombo);
    if (D == 24) {
      SLAPH_PROD_MFMA(24)
    } else {
      SLAPH_PROD_MFMA(32)
    }
#undef SLAPH_PROD_MFMA
    return;
  }
  if (D == 16 && !generic && !std::getenv("SLAPH_PROD_GENERIC")) {
    const int per_block = 4 * kProdPerWave;
    const dim3 grid((ncombo + per_block - 1) / per_block, ntime);
    if ((long long)grid.x * grid.y > 0x7fffffffLL) throw std::runtime_error("launch_prod: grid too large");
    if (transposed)
      hipLaunchKernelGGL((k_prod16<T, true>), grid, dim3(256), 0, s, x, Xt, Xq, y, Yt, Yq, o, obase, ostride, combos,
                         ncombo);
    else
      hipLaunchKernelGGL((k_prod16<T, false>), grid, dim3(256), 0, s, x, Xt, Xq, y, Yt, Yq, o, obase, ostride, combos,
                         ncombo);
  } else {
    hipLaunchKernelGGL(k_prod<T>, dim3(ncombo, ntime), dim3(64), 0, s, x, Xt, Xq, y, Yt, Yq, o, obase, ostride, combos,
                       D, transposed);
  }
}

template <typename T>
void launch_trace_lstat_t(const void *sX, const long long *Xt, const void *sY, const long long *Yt,
                          const LsBlock *blocks, int nblock, const LsChunk *chunks, const int *left,
                          const LsTask *tasks, const long long *slot_dst, double2 *out, long long obase,
                          long long ostride, int ntime, int D, int rt, hipStream_t s) {
  const T *x = static_cast<const T *>(sX), *y = static_cast<const T *>(sY);
  const dim3 grid(unsigned(nblock * ntime)), block(64 * kLsWaves);
  switch (D) {
    case 8:
      hipLaunchKernelGGL((k_trace_lstat<8, T>), grid, block, 0, s, x, Xt, y, Yt, blocks, nblock, chunks, left,
                         tasks, slot_dst, out, obase, ostride, rt);
      break;
    case 16:
      hipLaunchKernelGGL((k_trace_lstat<16, T>), grid, block, 0, s, x, Xt, y, Yt, blocks, nblock, chunks, left,
                         tasks, slot_dst, out, obase, ostride, rt);
      break;
    case 24:
      hipLaunchKernelGGL((k_trace_lstat<24, T>), grid, block, 0, s, x, Xt, y, Yt, blocks, nblock, chunks, left,
                         tasks, slot_dst, out, obase, ostride, rt);
      break;
    case 32:
      hipLaunchKernelGGL((k_trace_lstat<32, T>), grid, block, 0, s, x, Xt, y, Yt, blocks, nblock, chunks, left,
                         tasks, slot_dst, out, obase, ostride, rt);
      break;
    default:
      throw std::runtime_error("launch_trace_lstat: unsupported factor dimension");
  }
}
template <int D, typename T>
void launch_lstat_q_d(const T *x, const long long *Xt, const T *y, const long long *Yt, const int *units, int nunit,
                      const LsTask *tasks, const long long *slot_dst, double2 *out, long long obase, long long ostride,
                      int ntime, int rt, int *qcount, int wg_per_cu, hipStream_t s) {
  // wg_per_cu workgroups per compute unit (fewer if fewer fit), at most one
  // per unit of work.  With two the units in flight on one XCD span few
  // component-pieces and more of their right matrices stay in its L2: on the
  // 32^3 x 64 workload (ring of 4 at three waves per SIMD) 512 workgroups ran
  // the launch in 14.5 ms, 768 in 16.6 ms, 384 in 16.5 ms.  At two waves per
  // SIMD the registers allow a ring of 6: 14.2 ms (8: 14.8 ms).
  static int ncu = 0, per_cu = 0;
  if (ncu == 0) {
    int dev = 0;
    SLAPH_HIP(hipGetDevice(&dev));
    SLAPH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    SLAPH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_lstat_q<D, T>, 64 * kLsWaves, 0));
  }
  const int slots = std::max(1, ncu * std::max(1, std::min(per_cu, std::max(1, wg_per_cu))));
  const long long total = (long long)nunit * ntime;
  const unsigned grid = unsigned(std::min<long long>(total, slots));
  // descriptor window: the largest whole number of rounds that keeps a
  // round's reads inside the 64-lane batch; SLAPH_LSTAT_WINDOW=k selects k
  // rounds (tests use 1 to run the batch refresh at every round)
  const int pf = trace_lstat_qring(D);
  int win = pf * ((64 - pf) / pf);
  if (const char *e = std::getenv("SLAPH_LSTAT_WINDOW")) win = pf * std::max(1, std::min(std::atoi(e), (64 - pf) / pf));
  if (win % pf != 0 || win < pf || win > 64 - pf) throw std::logic_error("launch_trace_lstat_q: bad descriptor window");
  hipLaunchKernelGGL(k_ls_qreset, dim3(1), dim3(64), 0, s, qcount);
  hipLaunchKernelGGL((k_trace_lstat_q<D, T>), dim3(grid), dim3(64 * kLsWaves), 0, s, x, Xt, y, Yt, units, nunit, tasks,
                     slot_dst, out, obase, ostride, ntime, rt, qcount, win);
}

template <typename T>
void launch_trace_lstat_q_t(const void *sX, const long long *Xt, const void *sY, const long long *Yt, const int *units,
                            int nunit, const LsTask *tasks, const long long *slot_dst, double2 *out, long long obase,
                            long long ostride, int ntime, int D, int rt, int *qcount, int wg_per_cu,
                            hipStream_t s) {
  const T *x = static_cast<const T *>(sX), *y = static_cast<const T *>(sY);
  switch (D) {
    case 8:
      launch_lstat_q_d<8>(x, Xt, y, Yt, units, nunit, tasks, slot_dst, out, obase, ostride, ntime, rt, qcount,
                          wg_per_cu, s);
      break;
    case 16:
      launch_lstat_q_d<16>(x, Xt, y, Yt, units, nunit, tasks, slot_dst, out, obase, ostride, ntime, rt, qcount,
                          wg_per_cu, s);
      break;
    case 24:
      launch_lstat_q_d<24>(x, Xt, y, Yt, units, nunit, tasks, slot_dst, out, obase, ostride, ntime, rt, qcount,
                          wg_per_cu, s);
      break;
    case 32:
      launch_lstat_q_d<32>(x, Xt, y, Yt, units, nunit, tasks, slot_dst, out, obase, ostride, ntime, rt, qcount,
                          wg_per_cu, s);
      break;
    default:
      throw std::runtime_error("launch_trace_lstat_q: unsupported factor dimension");
  }
}
}  // namespace

void launch_trace_lstat_q(const void *sX, const long long *Xt, const void *sY, const long long *Yt, const int *units,
                          int nunit, const LsTask *tasks, const long long *slot_dst, double2 *out, long long obase,
                          long long ostride, int ntime, int D, int rt, int f32, int *qcount, int wg_per_cu,
                          hipStream_t s) {
  if (nunit <= 0 || ntime <= 0) return;
  if ((long long)nunit * ntime > 0x7fffffffLL) throw std::runtime_error("launch_trace_lstat_q: too many units");
  with_storage(f32, [&](auto tag) {
    launch_trace_lstat_q_t<decltype(tag)>(sX, Xt, sY, Yt, units, nunit, tasks, slot_dst, out, obase, ostride, ntime,
                                          D, rt, qcount, wg_per_cu, s);
  });
  SLAPH_HIP(hipGetLastError());
}

void launch_prod(const void *sX, const long long *Xt, const long long *Xq, const void *sY, const long long *Yt,
                 const long long *Yq, void *out, long long obase, long long ostride, const ComboProd *combos,
                 int ncombo, int ntime, int D, int transposed, int f32, int generic, hipStream_t s) {
  if (ncombo <= 0 || ntime <= 0) return;
  if (D > kMaxD) throw std::runtime_error("factor dimension above 32 is not supported");
  if (ntime > 65535) throw std::runtime_error("launch_prod: too many time keys");
  with_storage(f32, [&](auto tag) {
    launch_prod_t<decltype(tag)>(sX, Xt, Xq, sY, Yt, Yq, out, obase, ostride, combos, ncombo, ntime, D, transposed,
                                 generic, s);
  });
  SLAPH_HIP(hipGetLastError());
}

void launch_sprod(const void *sQ0, const long long *Xt, const void *sQ2, const long long *Yt, void *out,
                  long long obase, long long ostride, const STerm *terms, const SpUnit *units, int nunit, int ntime,
                  int f32, hipStream_t s) {
  if (nunit <= 0 || ntime <= 0) return;
  if (ntime > 65535) throw std::runtime_error("launch_sprod: too many time keys");
  if ((long long)nunit * ntime > 0x7fffffffLL) throw std::runtime_error("launch_sprod: grid too large");
  with_storage(f32, [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_sprod16<T>, dim3(nunit, ntime), dim3(64 * kSpWaves), 0, s, static_cast<const T *>(sQ0), Xt,
                       static_cast<const T *>(sQ2), Yt, static_cast<T *>(out), obase, ostride, terms, units, nunit);
  });
  SLAPH_HIP(hipGetLastError());
}

void launch_trace_sym(const void *slab, const long long *Tt, const SymUnit *units, int nunit, const int *stages,
                      const int *tiles, double2 *part, long long pstride, int ntime, int f32, hipStream_t s) {
  if (nunit <= 0 || ntime <= 0) return;
  if ((long long)nunit * ntime > 0x7fffffffLL) throw std::runtime_error("launch_trace_sym: grid too large");
  with_storage(f32, [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_trace_sym<T>, dim3(unsigned(nunit * ntime)), dim3(64 * kSymWaves), 0, s,
                       static_cast<const T *>(slab), Tt, units, ntime, stages, tiles, part, pstride);
  });
  SLAPH_HIP(hipGetLastError());
}

bool trace_lstat_supported(int D) { return D == 8 || D == 16 || D == 24 || D == 32; }

int trace_lstat_chunk(int D) {
  if (!trace_lstat_supported(D)) return 0;
  return kLsStageBytes / (D * D * int(sizeof(double2)));
}

int trace_lstat_ring(int D) {
  const int epl = D * D / 64;
  return epl <= 4 ? 4 : (epl <= 8 ? 2 : 1);
}

int trace_lstat_qring(int D) {
  const int epl = D * D / 64;
  return epl <= 4 ? 6 : (epl <= 8 ? 2 : 1);
}

int trace_lstat_qchunk(int D) {
  if (!trace_lstat_supported(D)) return 0;
  return std::min(trace_lstat_chunk(D), 64 - kLsUnitHdr - kLsWaves * trace_lstat_qring(D));
}

void launch_trace_lstat(const void *sX, const long long *Xt, const void *sY, const long long *Yt,
                        const LsBlock *blocks, int nblock, const LsChunk *chunks, const int *left,
                        const LsTask *tasks, const long long *slot_dst, double2 *out, long long obase,
                        long long ostride, int ntime, int D, int rt, int f32, hipStream_t s) {
  if (nblock <= 0 || ntime <= 0) return;
  if ((long long)nblock * ntime > 0x7fffffffLL) throw std::runtime_error("launch_trace_lstat: grid too large");
  with_storage(f32, [&](auto tag) {
    launch_trace_lstat_t<decltype(tag)>(sX, Xt, sY, Yt, blocks, nblock, chunks, left, tasks, slot_dst, out, obase,
                                        ostride, ntime, D, rt, s);
  });
  SLAPH_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_ls_reduce(const double2 *__restrict__ part, long long pstride,
                                                   const GroupItem *__restrict__ items, int nitems,
                                                   const int *__restrict__ pidx, double2 *__restrict__ out,
                                                   long long obase, long long ostride) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, ti = blockIdx.y;
  if (k >= nitems) return;
  const GroupItem it = items[k];
  const double2 *p = part + (long long)ti * pstride;
  double2 v = make_double2(0.0, 0.0);
  for (int c = 0; c < it.n; ++c) v = cadd(v, p[pidx[it.list_off + c]]);
  out[obase + (long long)ti * ostride + it.dst] = v;
}

void launch_ls_reduce(const double2 *part, long long pstride, const GroupItem *items, int nitems, const int *pidx,
                      double2 *out, long long obase, long long ostride, int ntime, hipStream_t s) {
  if (nitems <= 0 || ntime <= 0) return;
  if (ntime > 65535) throw std::runtime_error("launch_ls_reduce: too many slice pairs in one launch");
  hipLaunchKernelGGL(k_ls_reduce, dim3((nitems + 255) / 256, ntime), dim3(256), 0, s, part, pstride, items, nitems,
                     pidx, out, obase, ostride);
  SLAPH_HIP(hipGetLastError());
}

void launch_trace_groups(const void *sX, const long long *Xt, const long long *Xq, const void *sY,
                         const long long *Yt, const long long *Yq, const GroupItem *items, int nitems, const int2 *pairs,
                         double2 *out, long long obase, long long ostride, int ntime, int D, int rt, int f32,
                         hipStream_t s) {
  if (nitems <= 0 || ntime <= 0) return;
  if (ntime > 65535) throw std::runtime_error("launch_trace_groups: too many slice pairs");
  with_storage(f32, [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_trace_groups<T>, dim3(nitems, ntime), dim3(64), 0, s, static_cast<const T *>(sX), Xt, Xq,
                       static_cast<const T *>(sY), Yt, Yq, items, pairs, out, obase, ostride, D, rt);
  });
  SLAPH_HIP(hipGetLastError());
}

int trace_gemm_parts(int ntiles, int nrun, int wg_per_cu) {
  // At least wg_per_cu workgroups per compute unit, raised to a part count
  // that divides the runs evenly when one is near (parts of unequal length
  // leave workgroups waiting on the longest), at most 32 parts.
  static int ncu = 0;
  if (ncu == 0) {
    int dev = 0;
    SLAPH_HIP(hipGetDevice(&dev));
    SLAPH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  if (ntiles <= 0 || nrun <= 0) return 1;
  const int want = std::max(1, std::min({(std::max(1, wg_per_cu) * ncu + ntiles - 1) / ntiles, 32, nrun}));
  for (int kp = want; kp <= std::min({2 * want, 32, nrun}); ++kp)
    if (nrun % kp == 0) return kp;
  return want;
}

void launch_trace_gemm(const void *sX, const void *sY, const GridGTile *tiles, int ntiles, const int *runs, int nrun,
                       const GridTile *times, double2 *out, long long obase, long long ostride, int D, int rt, int f32,
                       int kparts, double2 *part, hipStream_t s) {
  if (ntiles <= 0) return;
  if (D * D % 16 != 0 || D > kMaxD) throw std::runtime_error("launch_trace_gemm: unsupported factor dimension");
  if (kparts < 1 || kparts > 65535 || (kparts > 1 && part == nullptr))
    throw std::runtime_error("launch_trace_gemm: bad K split");
  static_assert(kGemmT * 16 == 64, "the GEMM tile is 64 x 64");
  if ((long long)ntiles * kparts > 0x7fffffffLL) throw std::runtime_error("launch_trace_gemm: grid too large");
  const dim3 grid(ntiles, kparts);
  with_storage(f32, [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_trace_gemm<T>, grid, dim3(256), 0, s, static_cast<const T *>(sX), static_cast<const T *>(sY),
                       tiles, runs, nrun, times, out, obase, ostride, D, rt, part);
  });
  if (kparts > 1)
    hipLaunchKernelGGL(k_gemm_reduce, grid1d((long long)ntiles * 4096, 256), dim3(256), 0, s, part, kparts, tiles,
                       ntiles, times, out, obase, ostride);
  SLAPH_HIP(hipGetLastError());
}

void launch_trace_grid_r(const void *sX, const void *sY, const GridRGroup *groups, int ngroups, const int *runs,
                         const GridTile *tiles, int ntiles, double2 *out, long long obase, long long ostride, int D,
                         int rt, int f32, hipStream_t s) {
  if (ngroups <= 0 || ntiles <= 0) return;
  if (D % 4 != 0 || D > kMaxD) throw std::runtime_error("launch_trace_grid_r: unsupported factor dimension");
  if (ntiles > 65535) throw std::runtime_error("launch_trace_grid_r: too many tiles");
  with_storage(f32, [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_trace_grid_r<T>, dim3(ngroups, ntiles), dim3(256), 0, s, static_cast<const T *>(sX),
                       static_cast<const T *>(sY), groups, runs, tiles, out, obase, ostride, D, rt);
  });
  SLAPH_HIP(hipGetLastError());
}

void launch_trace_grid(const void *sX, const long long *Xq, const void *sY, const long long *Yq,
                       const GroupItem *items, int nitems, const int2 *pairs, const GridTile *tiles, int ntiles,
                       double2 *out, long long obase, long long ostride, int D, int rt, int f32, hipStream_t s) {
  if (nitems <= 0 || ntiles <= 0) return;
  if (D % 4 != 0 || D > kMaxD) throw std::runtime_error("launch_trace_grid: unsupported factor dimension");
  if (ntiles > 65535) throw std::runtime_error("launch_trace_grid: too many tiles");
  with_storage(f32, [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_trace_grid<T>, dim3(nitems, ntiles), dim3(256), 0, s, static_cast<const T *>(sX), Xq,
                       static_cast<const T *>(sY), Yq, items, pairs, tiles, out, obase, ostride, D, rt);
  });
  SLAPH_HIP(hipGetLastError());
}

void launch_to_storage(const double2 *in, void *out, long long n, int form, hipStream_t s) {
  if (n <= 0) return;
  if (form == 0) throw std::runtime_error("launch_to_storage: the FP64 form needs no conversion");
  with_storage(form, [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_to_storage<T>, grid1d(n, 256), dim3(256), 0, s, in, static_cast<T *>(out), n);
  });
  SLAPH_HIP(hipGetLastError());
}

void launch_finalize(const double2 *sum, int ncorr, int Lt, const int *orig, double2 *out, hipStream_t s) {
  const long long n = (long long)ncorr * Lt;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_finalize, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, sum, ncorr, Lt, orig, out);
  SLAPH_HIP(hipGetLastError());
}

void launch_zeta(const double2 *gv, long long gbase, long long gstride, const uint16_t *gm, const ZItem *items,
                 int nitems, const uint16_t *zmask, int nz, double2 *zv, long long zbase, long long zstride, int nsp,
                 hipStream_t s) {
  long long n = (long long)nsp * nitems;
  if (n <= 0 || nz <= 0) return;
  hipLaunchKernelGGL(k_zeta, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, gv, gbase, gstride, gm, items, nitems,
                     zmask, nz, zv, zbase, zstride, nsp);
  SLAPH_HIP(hipGetLastError());
}

void launch_resolve(const double2 *gv, const uint16_t *gm, const CorrJob *jobs, int ncorr, int nsp,
                    const double2 *zv, const uint16_t *zmask, int nz, double2 *cval, int kahan, int zonly,
                    hipStream_t s) {
  // slice pairs per thread: up to 8 while at least 2^18 threads remain
  int spt = 1;
  while (spt < 8 && (long long)ncorr * ((nsp + 2 * spt - 1) / (2 * spt)) >= (1LL << 18)) spt *= 2;
  long long n = (long long)ncorr * ((nsp + spt - 1) / spt);
  if (n <= 0) return;
  if ((n + 255) / 256 > 0x7fffffffLL) throw std::runtime_error("launch_resolve: grid too large");
  if (zonly && !kahan)
    hipLaunchKernelGGL(k_resolve_z, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, jobs, ncorr, nsp, zv, zmask, nz,
                       cval, spt);
  else
    hipLaunchKernelGGL(k_resolve, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, gv, gm, jobs, ncorr, nsp, zv,
                       zmask, nz, cval, kahan, spt);
  SLAPH_HIP(hipGetLastError());
}

void launch_accumulate(const double2 *cval, const int *sidx, const int4 *tgroups, int ngroups, int ncorr,
                       double2 *sum, hipStream_t s) {
  if (ncorr <= 0 || ngroups <= 0) return;
  hipLaunchKernelGGL(k_accumulate, grid1d((long long)ngroups * ncorr, 256), dim3(256), 0, s, cval, sidx, tgroups,
                     ngroups, ncorr, sum);
  SLAPH_HIP(hipGetLastError());
}

void launch_scale(double2 *x, long long n, double f, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale, grid1d(n, 256), dim3(256), 0, s, x, n, f);
  SLAPH_HIP(hipGetLastError());
}

void launch_fill_gauss(double2 *x, long long n, unsigned long long seed, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_gauss, grid1d(n, 256), dim3(256), 0, s, x, n, seed);
  SLAPH_HIP(hipGetLastError());
}

void launch_fill_z2z2(double2 *x, long long n, unsigned long long seed, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_z2z2, grid1d(n, 256), dim3(256), 0, s, x, n, seed);
  SLAPH_HIP(hipGetLastError());
}

void launch_gemm_nn_small(const double2 *A, const double2 *R, double2 *C, long long K, int N, hipStream_t s) {
  if (K <= 0 || N <= 0) return;
  dim3 grid(unsigned((K + 255) / 256), N);
  hipLaunchKernelGGL(k_gemm_nn_small, grid, dim3(256), 0, s, A, R, C, K, N);
  SLAPH_HIP(hipGetLastError());
}

}  // namespace slaph
