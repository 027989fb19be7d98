// Device kernels of the contraction path and their host launchers.
// All matrices are column-major complex<double> stored as double2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace slaph {

#define SLAPH_HIP(call)                                                                       \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " +   \
                               __FILE__ + ":" + std::to_string(__LINE__) + ": " #call);       \
  } while (0)

// One matrix product C = rowscale .* (A^H diag(phase[k/3]) B) of a batch.
// A: K x M (lda), B: K x N (ldb), C: M x N (ldc); phase/rowscale may be null.
struct AhBJob {
  const double2 *A;
  const double2 *B;
  double2 *C;
  const double2 *phase;
  const double2 *rowscale;
  long long lda, ldb, ldc;
};
void launch_ahb(const AhBJob *d_jobs, int njobs, int M, int N, int K, hipStream_t s);
// Allocation-free form: the job table lives in a fixed device array of
// kAhbNoallocJobs entries (64 KiB) per GPU.  ahb_noalloc_table(device) looks
// its address up once per device (hipGetSymbolAddress, with that device
// current) and returns the cached address afterwards.
constexpr int kAhbNoallocJobs = 1024;
constexpr int kMaxDevices = 64;
AhBJob *ahb_noalloc_table(int device);
// Grouped products of rpj-row jobs: group g stacks rows[job0 .. job0+njob)
// (row r of job j at rows[j].A + r*lda) against one right matrix B (K x N,
// ldb); job j's rows go to rows[j].C (rpj x N, ldc) scaled by rows[j].rowscale.
struct AhBRow {
  const double2 *A;
  const double2 *rowscale;
  double2 *C;
};
struct AhBGroup {
  const double2 *B;
  int job0, njob;
};
void launch_ahb_grouped(const AhBGroup *d_groups, int ngroups, int max_jobs, const AhBRow *d_rows, int rpj, int N,
                        int K, long long lda, long long ldb, long long ldc, hipStream_t s);

// VdaggerV of one timeslice for a group of up to kVdvGroup momenta:
// C[p] = V^H diag(phase[p][k/3]) V, V: K x N column-major (ld = K), C: N x N.
// phase[p] may be null (unit phase, e.g. for a Gram matrix).
constexpr int kVdvGroup = 6;
struct VdvJob {
  const double2 *V;
  const double2 *phase[kVdvGroup];
  double2 *C[kVdvGroup];
  int np;
  const double2 *W;     // right operand (same shape as V); nullptr: V itself
  const int *ex[2];     // Ozaki form: column exponents of V and of the right operand
};
// f32 != 0: single-pass FP32 MFMA (V and phases rounded to FP32, FP32
// accumulation), results widened to FP64.
void launch_vdv(const VdvJob *d_jobs, int njobs, int N, long long K, int f32, hipStream_t s);
// VdaggerV precision forms, one entry per job (np = 1; phase[0] may be
// null): form 2 = three-term split BF16 on the BF16 MFMA with FP32
// accumulation, form 3 = Ozaki int8 (kOzakiSlices slices of 7 bits, exact
// int32 slice products, FP64 combination; needs ex[0] / ex[1] from
// launch_colexp of V and of the right operand).
constexpr int kOzakiSlices = 7;
void launch_vdv_split(const VdvJob *d_jobs, int njobs, int N, long long K, int form, hipStream_t s);
// C = A^H B (N x N, column-major) for K x N column-major A and B with tall
// K, FP64 MFMA, K split into parts summed in order; `part` holds
// ahb_big_part_elems(njobs, N, K) complex numbers.
struct AhbBigJob {
  const double2 *A, *B;
  double2 *C;
};
int ahb_big_parts(int njobs, int N, long long K);
long long ahb_big_part_elems(int njobs, int N, long long K);
void launch_ahb_big(const AhbBigJob *d_jobs, int njobs, int N, long long K, double2 *part, hipStream_t s);
// e[n] with max_k |V[k, n]| < 2^e for each column of a K x N matrix
void launch_colexp(const double2 *V, long long K, int N, int *e, hipStream_t s);

// VdaggerV by a sitewise Fourier factorisation.  The momentum phase of site
// (x, y, z) factorises, exp(-i 2 pi (px x / Lx + py y / Ly + pz z / Lz)), so
//   C(p)[i,j] = sum_x ex(x) sum_y ey(y) sum_z ez(z) W_xyz[i,j],
//   W_xyz[i,j] = sum_c conj(V[3 site + c, i]) V[3 site + c, j].
// k_vdv_ft forms W of one site for a 16 x 16 (i, j) tile with three FP64
// MFMAs, transforms it over z on the vector ALUs (cos / sin sums, which give
// +pz and -pz together) and over y into per-(py, pz) sums U(x), stored per x.
// k_vdv_xsum sums U over x with the px phases into the N x N matrices.  Only
// tile pairs I <= J are formed: C(p)[j,i] = conj(C(-p)[i,j]) supplies the
// lower triangle.
// V(t) = V + t * vstride is read through FtLayout: element (k, n) (k = 3
// site + colour, eigenvector n) at k * kstride + n * nstride; columns n >=
// ncol of the last 16-column tile read column ncol - 1 (their outputs are
// never stored).  The engine reads a row-major copy made by launch_vrow,
// {Npad, 1, Npad}: reading the column-major input in place ({1, K, N}) costs
// 16 cache lines per load instead of 8 and was 11 ms slower per 32^3 x 64
// configuration than the copy (8 ms).
constexpr int kFtMaxC = 13;  // (py, pz) combinations of one pass
constexpr int kFtMaxK = 2;   // nonzero |pz| values of one pass
constexpr int kFtMaxLz = 128;  // largest Lz (z phases staged in LDS)
struct FtLayout {
  long long kstride, nstride;
  int ncol;
};
struct FtPass {
  int nc, nk;
  // Per combination, which z transform it reads: 0 = pz 0, 1 + 2s = +k_s,
  // 2 + 2s = -k_s, with k_s the s-th nonzero |pz| of the pass.
  int tsel[kFtMaxC];
};
// ey: nc x Ly phases exp(-i 2 pi py y / Ly); zcs: nk x Lz pairs
// (cos, sin)(2 pi k_s z / Lz).  U: [t][x][combination][tile pair][256].
void launch_vdv_ft(const double2 *V, long long vstride, FtLayout lay, int Lx, int Ly, int Lz, const int2 *tiles,
                   int ntp, int nt, FtPass const &pass, const double2 *ey, const double2 *zcs, double2 *U,
                   hipStream_t s);
// One output of the x sum: combination c of the pass, phases ex[k * Lx + x]
// = exp(-i 2 pi q x / Lx); lower == 0 stores C[t][j][i] (column-major
// element (i, j)) for every tile pair, lower == 1 stores conj at (j, i) for
// tile pairs I < J.  C points at timeslice 0; U holds timeslices t0 .. t0 + nt;
// tstride = N * N.
struct FtTarget {
  double2 *C;
  int c, lower;
};
void launch_vdv_xsum(const double2 *U, int nc, int ntp, int nt, int Lx, const int2 *tiles, const FtTarget *targets,
                     int ntarget, const double2 *ex, int N, int t0, long long tstride, hipStream_t s);
// Vr[t][k * Npad + n] = V[t][n * K + k] (zero for n >= N), nt timeslices.
void launch_vrow(const double2 *V, double2 *Vr, long long K, int N, int Npad, int nt, hipStream_t s);

// Q0 of one (t, operator, random pair): D x D.
// Q0 in groups of jobs that share VdV, rho2 and gamma: X = VdV rho2 (diluted)
// is formed once per group, then each job's D x D block from its rho1.
struct Q0Group {
  const double2 *vdv;   // N x N (daggered if needed)
  const double2 *rho2;  // random vector r2 at timeslice t: N x 4 (index 4v+d)
  int gamma;
  int job_off, njob;
};
struct Q0Out {
  const double2 *rho1;  // random vector r1 at timeslice t
  double2 *out;
};
void launch_q0(const Q0Group *d_groups, int ngroups, const Q0Out *d_outs, int N, int dilE, hipStream_t s);

// Q1 of one (t, b, operator, random pair): D x D.
struct Q1Job {
  const double2 *vdv;
  const double2 *rho1;
  const double2 *peram2;  // column-major perambulator of r2
  double2 *out;
  int t, b, gamma;
};
void launch_q1(const Q1Job *d_jobs, int njobs, int N, int dilE, long long peram_rows, hipStream_t s);

// Second stage of Q2: out = M' (D x 4N) * stacked perambulator rows.
struct Q2BJob {
  const double2 *Mp;      // D x 4N, column (bd*N + w)
  const double2 *peram2;
  double2 *out;
  int t, b2, gamma;
};
void launch_q2b(const Q2BJob *d_jobs, int njobs, int N, int dilE, long long peram_rows, hipStream_t s);

void launch_dagger(const double2 *in, double2 *out, int N, int count, hipStream_t s);
void launch_identity(double2 *out, int N, int count, hipStream_t s);
// n zero complex numbers by a kernel (a fill that a HIP graph may hold)
void launch_zero(double2 *p, long long n, hipStream_t s);

// Storage forms of factor slabs and product tables (the arithmetic is FP64 in
// every form): 0 = double2; 1 = float2 (FP32); 2 = d48c, each FP64 part cut to
// its upper 48 bits (sign, exponent and 36 of the 52 mantissa bits, rounded to
// nearest: relative rounding error at most 2^-37): w[0], w[1] = upper 32 bits
// of the real and imaginary part, w[2] = their next 16 bits (real in the low
// half).
struct __attribute__((aligned(4))) d48c {
  unsigned int w[3];
};
inline size_t storage_bytes(int form) {
  return form == 1 ? sizeof(float2) : (form == 2 ? sizeof(d48c) : sizeof(double2));
}

// Table-driven batched factor product / trace kernels.
// Operand matrix index = Xt[ti] + Xq[qn] + i; output index = obase + ti*ostride + o.
// transposed != 0 stores each product as (X Y)^T.  The operand slabs and the
// output are arrays of the storage form `f32` (0, 1 or 2 above; the name is
// kept from the two-form days).
struct ComboProd {
  int qn, i, j;
  long long o;
};
// generic != 0 (or SLAPH_PROD_GENERIC set): D = 16 products by the generic
// LDS kernel instead of the MFMA tiles
void launch_prod(const void *slabX, const long long *Xt, const long long *Xq, const void *slabY, const long long *Yt,
                 const long long *Yq, void *out, long long obase, long long ostride, const ComboProd *combos,
                 int ncombo, int ntime, int D, int transposed, int f32, int generic, hipStream_t s);
// Summed products of the crossed four-factor trace tr(Q0 Q2 Q0 Q2) (D = 16).
// A unit is (time key, one Q2 operator entry): its np pair matrices
// Q2[Yt + s0 .. + np) are staged in LDS, and its terms multiply one of them
// with a Q0 matrix:
//  kind 0: C = Q0[Xt + l] Q2[s0 + r], summed over the consecutive terms of
//          one output o (the last has bit 0 of flags set) and stored once;
//  kind 1: C = (Q2[s0 + r] Q0[Xt + l])^T = Q0^T Q2^T, stored at o; after the
//          last term of a run (bit 0) i times the run's sum is stored at o + 1;
// with Q0 / Q2 the factor slabs at the time bases Xt[ti] / Yt[ti], output
// matrix index obase + ti * ostride + o, column-major like launch_prod's.
// A Q0 factor is nonzero only in its Dirac blocks (row block perm(d), column
// block d) (DilutedFactorFactory.cpp:70-90): bits 8 + 2d of flags hold
// perm(d), the gamma matrix's row of column d; bit 1 is the kind.  Wave w of a
// unit runs terms [tb[w], tb[w + 1]), whole runs each.
constexpr int kSpStage = 20;  // largest np (80 KiB of LDS)
constexpr int kSpWaves = 8;   // wavefronts per unit
struct STerm {
  int l, r, o, flags;
};
struct SpUnit {
  int s0, np;
  int tb[kSpWaves + 1];
};
void launch_sprod(const void *slabQ0, const long long *Xt, const void *slabQ2, const long long *Yt, void *out,
                  long long obase, long long ostride, const STerm *terms, const SpUnit *units, int nunit, int ntime,
                  int f32, hipStream_t s);
// Traces of the summed-product tables (D = 16) as symmetric products.  The
// items of one connected component of a trace graph read the table matrices
// M_e(l) of their two entries (e1, e2) as
//   T[e1][e2] = sum_l tr(M_e1(l) M_e2(pi l)),
// pi a pairing of the labels (involution), so that T is symmetric.  A unit
// stages up to kSymEntries entries (positions p) and runs nst stages (an
// even number); stage k holds two 8 x 8 blocks (slot 0, slot 1) of a matrix
// of every staged entry and contributes two K steps of T = A B^T: "cross"
// (flags bit 0 clear) pairs
// slot 0 as A with slot 1 read transposed as B, then slot 1 with slot 0;
// "self" (bit 0 set) pairs each slot with its own transpose.  The stage table
// (stages + st_off) holds per stage 1 + 2 ns ints: the flags, then per
// position p the element offsets of its slot 0 and slot 1 blocks (first
// element) from the slice pair's time base, -1 for a zero block.  Tiles: per
// tile 17 ints from tiles + 17 tile_off: row group | column group << 8 (four
// positions each), then the partial-sum index of T[4 row + i][4 col + j] at
// 1 + 4 i + j, -1 where none is needed; wave w of a unit runs tiles
// [wt[w], wt[w + 1]) (at most kSymWaveTiles).  Output: part[ti * pstride +
// index] for the slice pairs ti < ntime, table base Tt[ti] (matrices).
constexpr int kSymEntries = 36;   // staged positions per unit (2 KiB of LDS each, two stage buffers)
constexpr int kSymWaves = 8;      // wavefronts per unit
constexpr int kSymWaveTiles = 6;  // tiles per wavefront
struct SymUnit {
  int st_off, nst, ns, tile_off;
  int wt[kSymWaves + 1];
};
void launch_trace_sym(const void *slab, const long long *Tt, const SymUnit *units, int nunit, const int *stages,
                      const int *tiles, double2 *part, long long pstride, int ntime, int f32, hipStream_t s);
// In the trace launchers below, rt != 0 means the right matrices are stored
// transposed (Y^T), so that tr(X Y) = sum_e X[e] Y^T[e]; the operand slabs
// are double2 or (f32 != 0) float2 arrays, the group sums always double2.
// FP64 values -> storage form `form` (1 or 2), n complex elements.
void launch_to_storage(const double2 *in, void *out, long long n, int form, hipStream_t s);

// Correlator resolution: mean / 2-way / 3-way masked inner products.
// A trace vector seen through its used-index mask groups: the trace values of
// all random-vector combinations with the same mask are summed first
// (group sums are exact re-associations of the reference's sums).
struct TraceRef {
  long long base;    // offset of the group sums at slice pair 0 (add s*stride)
  long long stride;  // group-sum stride per slice pair
  int ngroup;        // number of mask groups
  int gmask_off;     // into group masks[]
  long long zbase;   // offset of the subset sum Z(S_0) at slice pair 0, or -1
  long long zstride;  // subset-sum stride per slice pair
  int zq;             // stride between consecutive subsets S_q (the family's subset-sum items)
};
// Subset sums of one trace vector: Z(S) = sum of the group sums whose mask
// contains S, for every S of the subset list zmask[0..nz), stored subset-major
// per slice pair: Z(S_q) of item i at zbase + s * zstride + q * nitems + i.  The masked inner
// products then follow from Moebius inversion on the subset lattice,
// [m0 & m1 (& m2) == 0] = sum_{S subset of m0 & m1 (& m2)} (-1)^|S|, i.e.
// sum_S (-1)^|S| Z0(S) Z1(S) (Z2(S)).
struct ZItem {
  long long goff;  // the trace's group sums within one slice pair
  int ngroup, gmask_off;
};
void launch_zeta(const double2 *gvals, long long gbase, long long gstride, const uint16_t *gmasks, const ZItem *items,
                 int nitems, const uint16_t *zmask, int nz, double2 *zv, long long zbase, long long zstride, int nsp,
                 hipStream_t s);
struct CorrJob {
  int ntr;
  int out;   // the correlator's index in the output (jobs are ordered by their traces)
  double n;  // number of admitted combinations (host-computed, exact); the value is divided by it
  TraceRef tr[3];
};
// Trace group sums: out[obase + ti*ostride + dst] = sum over the item's
// combinations (i, j) of tr(X[Xt[ti] + Xq[qn] + i] * Y[Yt[ti] + Yq[qn] + j])
// (or tr(X[...]) when Y is null).
struct GroupItem {
  int qn;
  int list_off, n;
  long long dst;
};
void launch_trace_groups(const void *slabX, const long long *Xt, const long long *Xq, const void *slabY,
                         const long long *Yt, const long long *Yq, const GroupItem *items, int nitems, const int2 *pairs,
                         double2 *out, long long obase, long long ostride, int ntime, int D, int rt, int f32,
                         hipStream_t s);
// Time-separable traces: within one tile of slice pairs the left operand
// depends on one time index a < nL and the right one on b < nR, and every
// (a, b) is a slice pair sidx[a][b] (or -1).  Output per (item, tile):
// out[obase + sidx[a][b]*ostride + dst] = sum_combos tr(L_a,i R_b,j).
// An item's pairs must be sorted by their left index i.
constexpr int kGridT = 16;
struct GridTile {
  long long Lt[kGridT];
  long long Rt[kGridT];
  int sidx[kGridT][kGridT];
  int nL, nR;
};
void launch_trace_grid(const void *slabX, const long long *Xq, const void *slabY, const long long *Yq,
                       const GroupItem *items, int nitems, const int2 *pairs, const GridTile *tiles, int ntiles,
                       double2 *out, long long obase, long long ostride, int D, int rt, int f32, hipStream_t s);
// Right-stationary form of the same traces: a group is up to kGridIPW items
// that read one right operator entry (offset yq).  Its runs are records of
// kGridRun ints: the right combination(s) j1, j2 (j2 = -1: one matrix), then
// per item the left matrix offset (left entry + combination) whose trace pairs
// with R_j1 (+ R_j2), or -1.  Item w's tile goes to out[obase + sidx*ostride +
// dst[w]].  Every right pair is read once per group instead of once per item.
constexpr int kGridIPW = 4;
constexpr int kGridRun = 2 + kGridIPW;
struct GridRGroup {
  long long yq;
  int nitem, nrun, run_off, pad;
  long long dst[kGridIPW];
};
void launch_trace_grid_r(const void *slabX, const void *slabY, const GridRGroup *groups, int ngroups, const int *runs,
                         const GridTile *tiles, int ntiles, double2 *out, long long obase, long long ostride, int D,
                         int rt, int f32, hipStream_t s);
// GEMM form of the same traces when every item has the same runs (i_s, j1_s,
// j2_s): the tiles of all (left entry, right entry) pairs of a block of
// kGemmT x kGemmT entries are one complex GEMM, rows (left entry, left time),
// columns (right entry, right time), K = (run, matrix element):
//   C[(E, a)][(F, b)] = sum_s sum_e L(E, i_s, a)[e] (R(F, j1_s, b) + R(F, j2_s, b))[e'].
// dst[r][c] is the item of (xq[r], yq[c]) or -1; rows/columns past nE/nF and
// pairs without an item are computed and discarded.  runs: 3 ints per run.
constexpr int kGemmT = 4;
struct GridGTile {
  long long xq[kGemmT], yq[kGemmT];
  long long dst[kGemmT][kGemmT];
  int nE, nF;
};
// K (the runs) is split into kparts (trace_gemm_parts) parts; with more than
// one, part holds ntiles * kparts * 4096 double2 partial tiles that a second
// kernel sums in part order.
int trace_gemm_parts(int ntiles, int nrun, int wg_per_cu);
void launch_trace_gemm(const void *slabX, const void *slabY, const GridGTile *tiles, int ntiles, const int *runs,
                       int nrun, const GridTile *times, double2 *out, long long obase, long long ostride, int D, int rt,
                       int f32, int kparts, double2 *part, hipStream_t s);
// Non-separable traces, left-stationary.  A block is a set of (trace id,
// mask group) slots; its chunks stage left matrices X[Xt[ti] + left[loff + m]]
// (m < n, any entries and combinations) and list tasks per wavefront.  A task
// multiplies right matrix Y[Yt[ti] + yrel] with staged left matrices ia (and
// ib >= 0); slot >= 0 marks the last task of a run, whose sum is added to
// that slot.  Output: out[obase + ti*ostride + slot_dst[slot_off + k]] = slot k.
constexpr int kLsWaves = 4;
constexpr int kLsSlots = 64;
struct LsBlock {
  int chunk_off, nchunk;
  int slot_off, nslot;
};
struct LsChunk {
  int loff, n;
  int wbeg[kLsWaves], wcnt[kLsWaves];  // each wavefront's tasks [wbeg, wbeg + wcnt)
};
struct LsTask {
  int yrel;
  short ia, ib;
  int slot;
};
// Largest number of left matrices one chunk stages for matrix size D; each
// wavefront's task count per chunk must be a multiple of trace_lstat_ring(D),
// and a wavefront's tasks of one block are contiguous (chunk after chunk).
int trace_lstat_chunk(int D);
int trace_lstat_ring(int D);
bool trace_lstat_supported(int D);
void launch_trace_lstat(const void *slabX, const long long *Xt, const void *slabY, const long long *Yt,
                        const LsBlock *blocks, int nblock, const LsChunk *chunks, const int *left,
                        const LsTask *tasks, const long long *slot_dst, double2 *out, long long obase,
                        long long ostride, int ntime, int D, int rt, int f32, hipStream_t s);
// Queued form: every block has one chunk (at most trace_lstat_qchunk(D) left
// matrices), each wave's task count a multiple of trace_lstat_qring(D), given
// as a unit record of kLsUnitHdr ints (n, slot_off, nslot, wbeg[kLsWaves],
// wcnt[kLsWaves], 0), then kLsWaves * trace_lstat_qring(D) ints (each wave's
// first right matrices, 0 past its range), then
// trace_lstat_qchunk(D) left matrix offsets (the first n used).  Partial sum
// k of a unit goes to out[obase + ti*ostride + slot_off + k].  The task array
// ends with one null task.  qcount: 8 ints of device memory (one claim
// counter per XCD), cleared by the launch.
constexpr int kLsUnitHdr = 12;
int trace_lstat_qchunk(int D);
int trace_lstat_qring(int D);
// wg_per_cu: workgroups per compute unit of the persistent grid (at most
// what fits)
void launch_trace_lstat_q(const void *slabX, const long long *Xt, const void *slabY, const long long *Yt,
                          const int *units, int nunit, const LsTask *tasks, const long long *slot_dst, double2 *out,
                          long long obase, long long ostride, int ntime, int D, int rt, int f32, int *qcount,
                          int wg_per_cu, hipStream_t s);

// Launch shapes chosen by the autotuner (Engine::autotune); the defaults
// are the values measured best on the 32^3 x 64 workload.
struct LaunchTuning {
  int lstat_wg_per_cu = 2;  // persistent grid of the queued left-stationary traces
  int gemm_wg_per_cu = 16;  // workgroups per CU the K split of the GEMM-form traces aims at
  int prod_generic = 0;     // D = 16 products by the generic kernel instead of MFMA tiles
};
// Sums of partial slots: out[obase + ti*ostride + items[k].dst] =
// sum_{p in pidx[items[k].list_off .. + n)} part[ti*pstride + p], in list order.
void launch_ls_reduce(const double2 *part, long long pstride, const GroupItem *items, int nitems, const int *pidx,
                      double2 *out, long long obase, long long ostride, int ntime, hipStream_t s);
// kahan != 0: Kahan-Neumaier sums inside each (slice pair, correlator) in the
// reference's element order (every group one trace value, no subset sums).
// zonly != 0: every job has subset sums (tr[0].zbase >= 0, ntr > 1), so the
// Moebius form's own kernel may run.
void launch_resolve(const double2 *gvals, const uint16_t *gmasks, const CorrJob *jobs, int ncorr, int nsp,
                    const double2 *zv, const uint16_t *zmask, int nz, double2 *cval, int kahan, int zonly,
                    hipStream_t s);
// sum[t*ncorr + c] += cval[s*ncorr + c] for the slice pairs s of each time
// separation t, in the order listed: tgroups[g] = (t, offset into sidx,
// count), sidx ascending within a group.
void launch_accumulate(const double2 *cval, const int *sidx, const int4 *tgroups, int ngroups, int ncorr,
                       double2 *sum, hipStream_t s);
void launch_scale(double2 *x, long long n, double f, hipStream_t s);
// Correlator values (Diagram.cpp:132-147): out[o*Lt + t] = sum[t*ncorr + i] / Lt
// with o = orig[i] (the correlator's request index; i when orig is null).
void launch_finalize(const double2 *sum, int ncorr, int Lt, const int *orig, double2 *out, hipStream_t s);

// Synthetic inputs: seeded counter-based Gaussian / Z2xZ2 fills.
void launch_fill_gauss(double2 *x, long long n, unsigned long long seed, hipStream_t s);
void launch_fill_z2z2(double2 *x, long long n, unsigned long long seed, hipStream_t s);
// C (K x N) = A (K x N) * R (N x N), R column-major, all on device.
void launch_gemm_nn_small(const double2 *A, const double2 *R, double2 *C, long long K, int N, hipStream_t s);

}  // namespace slaph
