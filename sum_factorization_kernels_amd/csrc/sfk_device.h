// Device-side building blocks for the gfx950 sum-factorization kernels:
// even-odd 1D contraction tables (passed by value as kernel arguments, so
// every lane reads the same coefficient from scalar registers), the even-odd
// contraction of one register-resident line, and small index helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "sfk_internal.h"

namespace sfk {

// Even-odd form of the contraction out[o] = sum_k K[k][o] in[k] of an N x N
// matrix K with the centro-symmetry K[N-1-k][N-1-o] = SIG * K[k][o]
// (SIG = +1 for basis values, -1 for derivatives).  With e_k = in[k] +
// in[N-1-k], o_k = in[k] - in[N-1-k] (k < N/2):
//   P_o = sum_k A[k][o] e_k + C[o] in[mid],  Q_o = sum_k B[k][o] o_k,
//   out[o] = P_o + Q_o,  out[N-1-o] = SIG (P_o - Q_o),
// and for odd N the middle output sum_k E[k] (SIG>0 ? e_k : o_k) + Kmm in[mid].
// Halves the multiplications of the dense contraction (the reference kernel
// legacy/include/matrix_vector_kernel.h:30-185 uses the same symmetry; this
// form keeps the middle terms of every matrix, so it is exact also for
// bases that are not nodal at x = 1/2).
template <int N, typename T>
struct Eo {
  static constexpr int M = N / 2 > 0 ? N / 2 : 1;
  T A[M][M];
  T B[M][M];
  T C[M];
  T E[M];
  T Kmm;
};

// Build from a dense matrix M[i*n+q]; forward: K = M (out[q] = sum_i
// M[i][q] in[i]), backward: K = M^T (out[i] = sum_q M[i][q] in[q]).
template <int N, typename T>
inline Eo<N, T> make_eo(const std::vector<double> &Mat, bool forward)
{
  Eo<N, T> e{};
  auto K = [&](int k, int o) { return forward ? Mat[k * N + o] : Mat[o * N + k]; };
  const int mid = N / 2;
  for (int k = 0; k < mid; ++k)
    for (int o = 0; o < mid; ++o) {
      e.A[k][o] = T(0.5 * (K(k, o) + K(N - 1 - k, o)));
      e.B[k][o] = T(0.5 * (K(k, o) - K(N - 1 - k, o)));
    }
  if (N % 2 == 1) {
    for (int o = 0; o < mid; ++o) e.C[o] = T(K(mid, o));
    for (int k = 0; k < mid; ++k) e.E[k] = T(K(k, mid));
    e.Kmm = T(K(mid, mid));
  }
  return e;
}

template <int N, int SIG, typename E, typename T>
__device__ __forceinline__ void eo_apply(const E &t, const T (&in)[N], T (&out)[N])
{
  constexpr int mid = N / 2;
  T e[Eo<N, T>::M], o[Eo<N, T>::M];
#pragma unroll
  for (int k = 0; k < mid; ++k) {
    e[k] = in[k] + in[N - 1 - k];
    o[k] = in[k] - in[N - 1 - k];
  }
  T r[N];
#pragma unroll
  for (int q = 0; q < mid; ++q) {
    T p = t.A[0][q] * e[0];
    T s = t.B[0][q] * o[0];
#pragma unroll
    for (int k = 1; k < mid; ++k) {
      p = t.A[k][q] * e[k] + p;
      s = t.B[k][q] * o[k] + s;
    }
    if constexpr (N % 2 == 1) p = t.C[q] * in[mid] + p;
    r[q] = p + s;
    r[N - 1 - q] = SIG > 0 ? p - s : s - p;
  }
  if constexpr (N % 2 == 1) {
    T m = SIG > 0 ? t.Kmm * in[mid] : T(0);
#pragma unroll
    for (int k = 0; k < mid; ++k) m = t.E[k] * (SIG > 0 ? e[k] : o[k]) + m;
    r[mid] = m;
  }
#pragma unroll
  for (int q = 0; q < N; ++q) out[q] = r[q];
}

// 1D tables of one operator at compile-time size N and precision T.
template <int N, typename T>
struct Tab {
  Eo<N, T> Sf, Sb;  // basis values: forward (dof -> quad), backward (quad -> dof)
  Eo<N, T> Df, Db;  // collocation derivative at Gauss points, and its transpose
  Eo<N, T> If, Ib;  // Gauss Lagrange at GLL points (inverse of Sf), and transpose
  T fv0[N], fv1[N], fd0[N], fd1[N]; // l_q(0), l_q(1), l_q'(0), l_q'(1)
  T w[N];           // Gauss weights
  T hd;             // phi_0'(0) of the Hermite-like basis
};

// The tables of an operator live in device memory (one copy per handle) and
// the kernels read them through a constant-address-space pointer, so the
// coefficients of each sweep are scalar loads right before the sweep.  (As
// kernel arguments, or loaded once, they would all stay live in scalar
// registers for the whole kernel and spill.)  opaque_ptr hides the pointer
// value from the optimiser at every use site, which keeps it from hoisting
// all table loads to the kernel entry.
template <int N, typename T>
struct DevTab {
  Tab<N, T> t;
  T Dd[N][N]; // dense collocation derivative D[m][q] (tangential derivatives)
};

#define SFK_CONST_AS __attribute__((address_space(4)))

template <typename P>
__device__ __forceinline__ P opaque_ptr(P p)
{
  __asm__ volatile("" : "+s"(p));
  return p;
}

template <int N, typename T>
__device__ __forceinline__ const SFK_CONST_AS Tab<N, T> &ctab(const DevTab<N, T> *p)
{
  return ((const SFK_CONST_AS DevTab<N, T> *)opaque_ptr(p))->t;
}

template <int N, typename T>
__device__ __forceinline__ const SFK_CONST_AS DevTab<N, T> &cdevtab(const DevTab<N, T> *p)
{
  return *(const SFK_CONST_AS DevTab<N, T> *)opaque_ptr(p);
}

template <int N, typename T>
inline DevTab<N, T> make_devtab(const Tables1D &t)
{
  DevTab<N, T> d{};
  Tab<N, T> &tb = d.t;
  tb.Sf = make_eo<N, T>(t.S, true);
  tb.Sb = make_eo<N, T>(t.S, false);
  tb.Df = make_eo<N, T>(t.D, true);
  tb.Db = make_eo<N, T>(t.D, false);
  if ((int)t.Sinv.size() == N * N) {
    tb.If = make_eo<N, T>(t.Sinv, true);
    tb.Ib = make_eo<N, T>(t.Sinv, false);
  }
  for (int i = 0; i < N; ++i) {
    tb.fv0[i] = T(t.fv0[i]);
    tb.fv1[i] = T(t.fv1[i]);
    tb.fd0[i] = T(t.fd0[i]);
    tb.fd1[i] = T(t.fd1[i]);
    tb.w[i] = T(t.weights[i]);
    for (int q = 0; q < N; ++q) d.Dd[i][q] = T(t.D[i * N + q]);
  }
  tb.hd = T(t.hermite_d);
  return d;
}

// Line helpers inside one cell's LDS block (x fastest, row pitch NP).
template <int N, int NP, int DIR>
__device__ __forceinline__ int line_base(int l)
{
  if constexpr (DIR == 0) return l * NP;                      // l = k*N + j
  else if constexpr (DIR == 1) return (l / N) * N * NP + (l % N); // l = k*N + i
  else return (l / N) * NP + (l % N);                         // l = j*N + i
}
template <int N, int NP, int DIR>
constexpr int line_stride()
{
  return DIR == 0 ? 1 : (DIR == 1 ? NP : N * NP);
}

template <int N, int NP, int DIR, typename T>
__device__ __forceinline__ void lds_read_line(const T *cell, int l, T (&v)[N])
{
  const int b = line_base<N, NP, DIR>(l);
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = cell[b + i * line_stride<N, NP, DIR>()];
}
template <int N, int NP, int DIR, typename T>
__device__ __forceinline__ void lds_write_line(T *cell, int l, const T (&v)[N])
{
  const int b = line_base<N, NP, DIR>(l);
#pragma unroll
  for (int i = 0; i < N; ++i) cell[b + i * line_stride<N, NP, DIR>()] = v[i];
}

// Padded LDS block of a cell (x fastest, rows of N, planes PP values apart,
// cells CS apart).  LDS serves 8-byte accesses in passes of 16 lanes and
// 16-byte accesses in passes of 8 lanes over 32 four-byte banks (measured on
// MI355X with the bank-conflict counter).  For N = 4 in FP64 an unpadded
// block (PP = 16) puts the 16 y lines of a cell on 4 bank pairs (4-way
// conflicts) and the x lines' 16-byte accesses 2-way; PP = 18 spreads the x
// lines over all 16 bank pairs and leaves the y lines 2-way.  For N = 4 in
// FP32, PP = 20 with cells 80 values apart measured fastest of the layouts
// tried (precomputed DG p = 3: 29 -> 25 us; planes 36 apart: 28 us).  Other
// N keep the unpadded block.
template <int N, typename T>
struct CellLds {
  static constexpr bool pad = N == 4;
  static constexpr int PP = pad ? (sizeof(T) == 8 ? 18 : 20) : N * N;
  static constexpr int CS = pad ? (sizeof(T) == 8 ? 3 * 18 + 16 : 80) : N * N * N;
};
// LDS offset of dof d (x fastest) inside a padded cell block
template <int N, typename T>
__device__ __forceinline__ int cl_off(int d)
{
  return (d / (N * N)) * CellLds<N, T>::PP + d % (N * N);
}
// line l of direction DIR of a padded cell block (x line l = k N + j, y line
// l = k N + i, z line l = j N + i)
template <int N, int DIR, typename T>
__device__ __forceinline__ void cl_read(const T *cell, int l, T (&v)[N])
{
  constexpr int PP = CellLds<N, T>::PP;
  const int b = DIR == 0 ? (l / N) * PP + (l % N) * N : (DIR == 1 ? (l / N) * PP + l % N : l);
  constexpr int st = DIR == 0 ? 1 : (DIR == 1 ? N : PP);
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = cell[b + i * st];
}
template <int N, int DIR, typename T>
__device__ __forceinline__ void cl_write(T *cell, int l, const T (&v)[N])
{
  constexpr int PP = CellLds<N, T>::PP;
  const int b = DIR == 0 ? (l / N) * PP + (l % N) * N : (DIR == 1 ? (l / N) * PP + l % N : l);
  constexpr int st = DIR == 0 ? 1 : (DIR == 1 ? N : PP);
#pragma unroll
  for (int i = 0; i < N; ++i) cell[b + i * st] = v[i];
}

// quadrature weight of point i on line l of direction DIR (w: the device
// table; the per-lane entries are vector loads)
template <int N, int DIR, typename P>
__device__ __forceinline__ auto line_weight(P w, int l, int i)
{
  return w[i] * (w[l % N] * w[l / N]);
}

// dof index (inside a cell) of point `layer` along DIR at face point fa
template <int N, int DIR>
__device__ __forceinline__ int face_dof(int layer, int fa)
{
  if constexpr (DIR == 0) return fa * N + layer;
  else if constexpr (DIR == 1) return (fa / N) * N * N + layer * N + (fa % N);
  else return layer * N * N + fa;
}

// the six periodic neighbours of cell ci (x fastest): below / above in x,
// then y, then z (cell counts fit in 32 bits)
__device__ __forceinline__ void periodic_neighbors(int ci, int nx, int ny, int nz, int (&nbr)[6])
{
  const int ix = ci % nx, r = ci / nx, iy = r % ny, iz = r / ny;
  const int sy = nx, sz = nx * ny;
  nbr[0] = ci + (ix == 0 ? nx - 1 : -1);
  nbr[1] = ci + (ix == nx - 1 ? 1 - nx : 1);
  nbr[2] = ci + (iy == 0 ? (ny - 1) * sy : -sy);
  nbr[3] = ci + (iy == ny - 1 ? -(ny - 1) * sy : sy);
  nbr[4] = ci + (iz == 0 ? (nz - 1) * sz : -sz);
  nbr[5] = ci + (iz == nz - 1 ? -(nz - 1) * sz : sz);
}

// Workgroups of a launch are dispatched round-robin over the chip's 8 XCDs
// (workgroup b starts on XCD b % 8), and every XCD has its own L2.  Block
// order 1 hands XCD x the contiguous range [start(x), start(x+1)) of
// logical blocks, so neighbouring cells, whose face layers and face metric
// each block re-reads, are processed by one XCD and meet in its L2.  Any
// order is a bijection of [0, nb): results do not depend on it.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned logical_block(unsigned b, unsigned nb, int order)
{
  if (order == 0) return b;
  const unsigned q = nb / kXcds, r = nb % kXcds, x = b % kXcds, k = b / kXcds;
  return x * q + (x < r ? x : r) + k;
}

// Cell at position q of a DG kernel's traversal.  Block order 2 walks slabs
// of kSlab z-layers x fastest, then z, then y: a cell's lower z and y
// neighbours (whose face layers and face metric it re-reads) were processed
// nx and nx*kSlab positions earlier, by the same XCD, and are still in its
// L2 (in index order the lower z neighbour is nx*ny positions back).
// Orders 0 and 1 walk the cells in index order.  A bijection of [0, ncells).
constexpr int kSlab = 4;
__device__ __forceinline__ long long traverse_cell(long long q, int nx, int ny, int nz, int order)
{
  if (order != 2) return q;
  // cell counts fit in 31 bits: 32-bit unsigned arithmetic throughout
  const unsigned uq = (unsigned)q, unx = (unsigned)nx, layer = unx * (unsigned)ny;
  const unsigned z0 = uq / layer / kSlab * kSlab;
  const unsigned zs = (unsigned)nz - z0 < (unsigned)kSlab ? (unsigned)nz - z0 : (unsigned)kSlab;
  const unsigned r = uq - z0 * layer;
  const unsigned r2 = r / unx;
  const unsigned x = r - r2 * unx, y = r2 / zs, zz = r2 - y * zs;
  return (long long)(((z0 + zz) * (unsigned)ny + y) * unx + x);
}

// Coordinates (x, y, z) of the cell at traversal position q (traverse_cell's
// order; cell = (z ny + y) nx + x).
struct CellXYZ {
  unsigned x, y, z;
};
__device__ __forceinline__ CellXYZ traverse_xyz(unsigned q, unsigned nx, unsigned ny, unsigned nz, int order)
{
  CellXYZ p;
  if (order != 2) {
    const unsigned r = q / nx;
    p.x = q - r * nx;
    p.y = r % ny;
    p.z = r / ny;
  } else {
    const unsigned layer = nx * ny;
    const unsigned z0 = q / layer / kSlab * kSlab;
    const unsigned zs = nz - z0 < (unsigned)kSlab ? nz - z0 : (unsigned)kSlab;
    const unsigned r = q - z0 * layer, r2 = r / nx;
    p.x = r - r2 * nx;
    p.y = r2 / zs;
    p.z = z0 + (r2 - p.y * zs);
  }
  return p;
}

// Cell c of a workgroup whose first position q0 is the same for all its
// threads: q0's coordinates are uniform (scalar unit), and position q0 + c
// of both traversals is q0's cell moved c steps in x unless that crosses
// the end of the x row (then the divisions are done per thread).  Returns
// the cell index; its periodic face neighbours go to nbr (as
// periodic_neighbors).
__device__ __forceinline__ long long locate_cell(unsigned q0, unsigned c, int nx, int ny, int nz, int order,
                                                 int (&nbr)[6])
{
  const unsigned unx = (unsigned)nx, uny = (unsigned)ny, unz = (unsigned)nz;
  CellXYZ p = traverse_xyz(q0, unx, uny, unz, order);
  if (p.x + c < unx) p.x += c;
  else p = traverse_xyz(q0 + c, unx, uny, unz, order);
  const int sy = nx, sz = nx * ny;
  const int ci = ((int)p.z * ny + (int)p.y) * nx + (int)p.x;
  nbr[0] = ci + (p.x == 0 ? nx - 1 : -1);
  nbr[1] = ci + (p.x == unx - 1 ? 1 - nx : 1);
  nbr[2] = ci + (p.y == 0 ? (ny - 1) * sy : -sy);
  nbr[3] = ci + (p.y == uny - 1 ? -(ny - 1) * sy : sy);
  nbr[4] = ci + (p.z == 0 ? (nz - 1) * sz : -sz);
  nbr[5] = ci + (p.z == unz - 1 ? -(nz - 1) * sz : sz);
  return ci;
}

// Waves per SIMD that a kernel's LDS allows at its default cells per
// workgroup (gfx950: 160 KiB of LDS per CU, allocated to workgroups in
// 1280-byte granules -- measured on MI355X with fixed-duration workgroups:
// 32000 B per workgroup still fits five per CU, 32256 B only four -- 4
// SIMDs, at most 8 waves per SIMD).  The operator kernels size their LDS at
// run time (dynamic LDS), so the compiler cannot see that LDS caps their
// occupancy; passing this bound as amdgpu_waves_per_eu(1, W) lets it spend
// up to 512/W registers per lane instead of squeezing them for waves that
// can never be resident.
constexpr int kLdsGranule = 1280;
constexpr int lds_waves_per_simd(int lds_bytes_per_block, int threads_per_block)
{
  const int alloc = (lds_bytes_per_block + kLdsGranule - 1) / kLdsGranule * kLdsGranule;
  const int blocks = alloc > 0 ? 163840 / alloc : 32;
  const int waves = blocks * ((threads_per_block + 63) / 64);
  const int w = (waves + 3) / 4;
  return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// read-once streams (stored metric): non-temporal loads keep them from
// evicting the re-read neighbour data out of L2
template <typename T>
__device__ __forceinline__ T load_stream(const T *p)
{
  return __builtin_nontemporal_load(p);
}

// dynamic LDS of a kernel (sized by the launch: cells per block is a
// run-time schedule parameter)
template <typename T>
__device__ __forceinline__ T *dyn_lds()
{
  extern __shared__ __attribute__((aligned(16))) unsigned char sfk_dyn_lds[];
  return reinterpret_cast<T *>(sfk_dyn_lds);
}

} // namespace sfk
