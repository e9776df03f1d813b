// Continuous Q_p Laplacian on an affine Cartesian mesh with a constant
// coefficient, evaluated as a sum of Kronecker products of the assembled
// 1D element matrices (tensor-product form).
//
// With Gauss quadrature of p+1 points per direction the element mass and
// stiffness matrices of the GLL-Lagrange basis are integrated exactly, so
// the cell operator of legacy/include/evaluation_cell_laplacian.h:264-503
// (value sweeps, collocation gradient, q-point operator G = diag(JxW J_d^2),
// transposed sweeps) equals
//   A_cell = c_x K (x) M (x) M + c_y M (x) K (x) M + c_z M (x) M (x) K,
//   M_ij = sum_q w_q phi_i(x_q) phi_j(x_q),  K_ij = sum_q w_q phi_i'(x_q) phi_j'(x_q),
//   c_d = JxW J_d^2 (the same constants as the cell kernel's cdiag),
// and on a structured grid of cells the assembled operator is the same sum
// with the assembled (banded) 1D matrices M^G, K^G in place of M, K.  The
// kernel applies it direction by direction on the global node grid,
//   a = M_y u,  b = c_y K_y u;   c = M_x a,  d = M_x b + c_x K_x a;
//   A u = M_z d + c_z K_z c,
// so nodes shared by cells are summed inside each 1D sweep: no scatter,
// colouring or atomics.  Every 1D sweep is the even-odd contraction of an
// element line (eo_apply) plus, for the element's first node (shared with
// the element before it), the last row of the element matrix applied to the
// element before: outputs of a line are its nodes 0..p-1.  Results equal the
// quadrature form up to rounding (the test suite pins both against the
// oracle at 1e-12).  Dirichlet nodes are 0 on input and output.
//
// A workgroup owns a tile of TX x TY nodes (BX x BY cells) in x/y and marches
// through a range of cells in z:
//   phase Y  (per z layer, the p new node planes): y lines of the input read
//            from HBM (lanes on consecutive x: coalesced), for the tile's rows
//            and a halo of p+1 columns (the x sweep needs the cell before the
//            tile and the tile's last node) -> a, b in LDS;
//   phase X  x lines of a, b from LDS -> c, d in LDS;
//   phase Z  one thread per node column: z line of c, d (the layer's first
//            plane kept in registers from the layer before), z sweep, stores
//            of the p finished planes (coalesced); the sum for the shared
//            plane is carried in a register to the next layer.
// Two barriers per layer.  A z range that starts above z = 0 first rebuilds
// the carried state from the cell below (one plane, then one layer without
// stores).  Optionally the Chebyshev update of do_cheby_iter<true>
// (evaluation_dg_laplacian.h:485-502) is fused into phase Z: the assembled
// value of a node is final there, and all its streams are coalesced.
#include <type_traits>

#include "sfk_device.h"
#include "sfk_kernels.h"

#ifndef SFK_TP_SCHED
#define SFK_TP_SCHED 1
#endif
#ifndef SFK_TP_WAVES
#define SFK_TP_WAVES 1
#endif

namespace sfk {

// tables of the tensor-product operator (device memory, read through the
// constant address space like DevTab)
template <int N, typename T>
struct TpTab {
  Eo<N, T> M;    // element mass
  Eo<N, T> K[3]; // c_d K, d = x, y, z
  T mrow[N];     // last row of M
  T krow[3][N];  // last row of c_d K
  // diagonals of the assembled 1D matrices at local node index i (i = 0: a
  // vertex shared by two cells): the operator's diagonal at node (i, j, k)
  // is dk[0][i] dm[j] dm[k] + dm[i] dk[1][j] dm[k] + dm[i] dm[j] dk[2][k]
  T dm[N];
  T dk[3][N];
};

template <int N, typename T>
__device__ __forceinline__ const SFK_CONST_AS TpTab<N, T> &tptab(const void *p)
{
  return *(const SFK_CONST_AS TpTab<N, T> *)opaque_ptr(p);
}

constexpr int odd_up(int v) { return v | 1; }
constexpr int round64(int v) { return (v + 63) / 64 * 64; }
constexpr int imax(int a, int b) { return a > b ? a : b; }

template <int N, typename T, int BX, int BY>
struct TpCfg {
  static constexpr int p = N - 1;
  static constexpr int TX = BX * p, TY = BY * p;
  static constexpr int HX = TX + p + 1;   // columns of phase Y (halo of p before, 1 after)
  static constexpr int RPA = odd_up(HX);  // row pitches: odd, so that 16 lanes on
  static constexpr int RPC = odd_up(TX);  // consecutive rows hit distinct banks
  static constexpr int NYI = p * BY * HX; // work items of phase Y (p planes)
  static constexpr int NXI = p * TY * BX; // phase X
  static constexpr int NZI = TX * TY;     // phase Z (node columns)
  static constexpr int THREADS_RAW = round64(imax(NYI, imax(NXI, NZI)));
  static constexpr int THREADS = THREADS_RAW > 1024 ? 1024 : THREADS_RAW;
  static constexpr int LA = p * TY * RPA; // one of a, b
  static constexpr int LC = p * TY * RPC; // one of c, d
  static constexpr int LDS_VALUES = 2 * LA + 2 * LC;
};

// waves per SIMD that the LDS of one tile shape allows (160 KiB per CU in
// 1280-byte granules, at most 32 waves per CU): the kernel asks the
// compiler for registers that let this many waves be resident, since an
// allocation above 512/W VGPRs would drop a whole workgroup per CU
template <int N, typename T, int BX, int BY>
constexpr int tp_waves()
{
  using C = TpCfg<N, T, BX, BY>;
  constexpr int bytes = C::LDS_VALUES * (int)sizeof(T);
  constexpr int alloc = (bytes + 1279) / 1280 * 1280;
  constexpr int wpg = C::THREADS / 64;
  int wgs = 163840 / alloc;
  if (wgs * wpg > 32) wgs = 32 / wpg;
  const int w = (wgs * wpg + 3) / 4;
  return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// waves per SIMD the kernel is compiled for: the matvec (CHEBY 0) asks for
// what the LDS allows; the Chebyshev forms hold more streams in registers
// (prefetched rhs, diagonal, tm / previous iterate) and would spill at that
// occupancy, so unless SFK_TP_CHEBY_WAVES is set they ask for one wave only
// (the register count then sets the occupancy)
#ifndef SFK_TP_CHEBY_WAVES
#define SFK_TP_CHEBY_WAVES 0
#endif
template <int N, typename T, int BX, int BY, int CHEBY>
constexpr int tp_waves_of()
{
  return CHEBY == 0 || CHEBY == 5 ? tp_waves<N, T, BX, BY>()
                    : (SFK_TP_CHEBY_WAVES > 0 ? SFK_TP_CHEBY_WAVES
                                              : (SFK_TP_CHEBY_WAVES < 0 ? tp_waves<N, T, BX, BY>() : 1));
}

// element at a 32-bit byte offset (per lane) from a uniform base pointer:
// the load or store takes the scalar base and the vector offset directly
// (no per-lane 64-bit address arithmetic)
template <typename T>
__device__ __forceinline__ const T &at(const T *base, unsigned off)
{
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T>
__device__ __forceinline__ T &at(T *base, unsigned off)
{
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off);
}

template <typename T>
struct TpArgs {
  const void *tab;
  const T *src;
  T *dst;
  const T *rhs, *diag; // fused Chebyshev
  T *tm;
  const T *prev;       // three-term Chebyshev: the previous iterate (or null: zero)
  T c_np, c_tm;
  // three-term Chebyshev, CHEBY 3 / 4: the input / the previous iterate is
  // the first iterate s_rhs D^{-1} rhs, formed in the kernel from rhs
  // instead of read (CHEBY 3: src points to rhs)
  T s_rhs;
  double *dot_part;  // CHEBY 5: per-workgroup partial sums of src . dst
  int nx, ny, nz;    // cells
  int gx, gy, gz;    // nodes
  int ntx, nty;      // tiles per direction
  int zc;            // cells in z per workgroup
  int order;         // block order (logical_block)
  long long vstride; // elements between right-hand sides
};

// element line with the shared first node: in[0..2p] = values at nodes
// e p - p .. e p + p; out[i] = (A u_e)[i] + (i == 0 ? (A u_{e-1})[p] : 0),
// i < p, for A = M (W = 0) or c_d K (W = 1 + d).  The table pointer is
// re-derived here and a scheduling barrier closes the contraction, so each
// matrix lives in scalar registers only while it is used (held for a whole
// phase, the tables of two matrices and their rows exceed the SGPR file).
template <int N, typename T, int W>
__device__ __forceinline__ void line_one(const void *tab, const T (&in)[2 * N - 1], T (&o)[N - 1])
{
  constexpr int p = N - 1;
  const auto &tb = tptab<N, T>(tab);
  const auto &E = W == 0 ? tb.M : tb.K[W > 0 ? W - 1 : 0];
  const auto &R = W == 0 ? tb.mrow : tb.krow[W > 0 ? W - 1 : 0];
  T ue[N], r[N];
#pragma unroll
  for (int i = 0; i < N; ++i) ue[i] = in[p + i];
  eo_apply<N, 1>(E, ue, r);
  T v = R[0] * in[0];
#pragma unroll
  for (int j = 1; j < N; ++j) v = R[j] * in[j] + v;
  o[0] = r[0] + v;
#pragma unroll
  for (int i = 1; i < p; ++i) o[i] = r[i];
#if SFK_TP_SCHED
  __builtin_amdgcn_sched_barrier(0);
#endif
}

// CHEBY: 0 matvec; 1 the Chebyshev step of do_cheby_iter<true>
// (evaluation_dg_laplacian.h:485-502): tm = c_tm tm + c_np diag (A u - rhs),
// dst = u - tm; 2 the same recurrence in three-term form with the
// operator's own inverse diagonal (solver smoother): with tm = prev - u,
// dst = u - c_tm (prev - u) - c_np D^{-1} (A u - rhs), four streams instead
// of six, and dst may be prev (each node is read, then written, by one
// thread); 3 and 4 the three-term step whose input (3) or previous iterate
// (4) is the first iterate s_rhs D^{-1} rhs, formed from rhs in the kernel
template <int N, typename T, int BX, int BY, int CHEBY>
__global__ void __launch_bounds__((TpCfg<N, T, BX, BY>::THREADS))
#if SFK_TP_WAVES
__attribute__((amdgpu_waves_per_eu(tp_waves_of<N, T, BX, BY, CHEBY>(), 8)))
#endif
cont_tp_kernel(const TpArgs<T> a)
{
  using C = TpCfg<N, T, BX, BY>;
  constexpr int p = N - 1, TX = C::TX, TY = C::TY, HX = C::HX, RPA = C::RPA, RPC = C::RPC;
  // three-term form (2, 3, 4); 3: the input, 4: the previous iterate is
  // s_rhs D^{-1} rhs (the first iterate, never stored)
  constexpr bool TT = CHEBY >= 2 && CHEBY <= 4, RSRC = CHEBY == 3, RPREV = CHEBY == 4, DOT = CHEBY == 5;
  // CHEBY 5: a layer's outputs held in registers and stored, with the
  // prefetch of the next layer's input at the node, after the next layer's
  // wait for its y lines (see flush below)
  constexpr bool DELAY = DOT;
  T *sA = dyn_lds<T>();
  T *sB = sA + C::LA;
  T *sC = sB + C::LA;
  T *sD = sC + C::LC;

  const unsigned L = logical_block(blockIdx.x, gridDim.x, a.order);
  const int tx = (int)(L % (unsigned)a.ntx);
  const unsigned r0 = L / (unsigned)a.ntx;
  const int ty = (int)(r0 % (unsigned)a.nty);
  const int chunk = (int)(r0 / (unsigned)a.nty);
  const int ex0 = tx * BX, ey0 = ty * BY;
  const int x0 = ex0 * p, y0 = ey0 * p;
  const int ez_begin = chunk * a.zc;
  const int ez_end = ez_begin + a.zc < a.nz ? ez_begin + a.zc : a.nz;
  const int gx = a.gx, gy = a.gy, gz = a.gz;
  const long long plane = (long long)gx * gy;
  const long long rr = blockIdx.y;
  const T *src = a.src + rr * a.vstride;
  T *dst = a.dst + rr * a.vstride;

  // phase Y, split in two so that the loads of the next layer are in flight
  // during phases X and Z of the current one: load_y fetches this thread's
  // y line (planes Z = ez p + k0 + s, s < NK, one work item per thread:
  // NK BY HX <= THREADS) into registers, store_y contracts it into slot s.
  static_assert(p * BY * HX <= C::THREADS, "one phase-Y item per thread");
  // the y lines of the tile's cells ey0 .. ey0+BY-1 span the rows
  // (ey0-1) p .. (ey0+BY) p: all interior node rows unless the tile is
  // within one cell of the y boundary
  const bool ybnd = !(ey0 >= 2 && ey0 + BY < a.ny);
  T u[2 * N - 1];
  bool yact = false;
  auto load_y_to = [&](T(&uu)[2 * N - 1], bool &act, int ez, int k0, auto nkc) {
    constexpr int NK = decltype(nkc)::value;
    const int it = threadIdx.x;
    const int xh = it % HX, rest = it / HX, eyl = rest % BY, s = rest / BY;
    const int ey = ey0 + eyl;
    // rows beyond the mesh are never read
    act = it < NK * BY * HX && ey < a.ny;
    const int X = x0 - p + xh, Z = ez * p + k0 + s;
    if (act && X > 0 && X < gx - 1 && Z > 0 && Z < gz - 1) {
      if (!ybnd) {
        // interior tile row range: every row of the line is an interior
        // node row; row j of all lanes is one uniform base + the lane's
        // 32-bit offset within the layer (the host checks that the layer's
        // planes span less than 2 GiB)
        const T *lb = src + (long long)(ez * p) * plane;
        const unsigned off = (unsigned)(((k0 + s) * (int)plane + (ey * p - p) * gx + X) * (int)sizeof(T));
#pragma unroll
        for (int j = 0; j < 2 * N - 1; ++j) uu[j] = at(lb + j * gx, off);
      } else {
        const T *col = src + (long long)Z * plane + X;
#pragma unroll
        for (int j = 0; j < 2 * N - 1; ++j) {
          const int Y = ey * p - p + j;
          uu[j] = (Y > 0 && Y < gy - 1) ? col[(long long)Y * gx] : T(0);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 2 * N - 1; ++j) uu[j] = T(0);
    }
    if constexpr (RSRC) {
      {
        // the line holds rhs: scale node j by s_rhs / D, D = a_xz dm[j % p] +
        // b_xz dk_y[j % p] with the x/z factors of this line (row j of a
        // line is node j % p of its cell; index 0 is the shared vertex);
        // the Dirichlet nodes were loaded as 0
        const auto &tb = tptab<N, T>(a.tab);
        const int ix = xh % p, iz = (k0 + s) % p;
        const T axz = tb.dk[0][ix] * tb.dm[iz] + tb.dm[ix] * tb.dk[2][iz];
        const T bxz = tb.dm[ix] * tb.dm[iz];
#pragma unroll
        for (int j = 0; j < 2 * N - 1; ++j) uu[j] *= a.s_rhs / (axz * tb.dm[j % p] + bxz * tb.dk[1][j % p]);
      }
    }
  };
  auto load_y = [&](int ez, int k0, auto nkc) { load_y_to(u, yact, ez, k0, nkc); };
  auto store_y_from = [&](const T(&uu)[2 * N - 1], bool act) {
    if (!act) return;
    const int it = threadIdx.x;
    const int xh = it % HX, rest = it / HX, eyl = rest % BY, s = rest / BY;
    T va[N - 1], vb[N - 1];
    line_one<N, T, 0>(a.tab, uu, va); // a = M_y u
    line_one<N, T, 2>(a.tab, uu, vb); // b = c_y K_y u
    T *ra = sA + (s * TY + eyl * p) * RPA + xh;
    T *rb = sB + (s * TY + eyl * p) * RPA + xh;
#pragma unroll
    for (int i = 0; i < p; ++i) {
      ra[i * RPA] = va[i];
      rb[i * RPA] = vb[i];
    }
  };
  auto store_y = [&]() { store_y_from(u, yact); };

  // phase X: rows of slots s < NK
  auto phase_x = [&](auto nkc) {
    constexpr int NK = decltype(nkc)::value;
    for (int it = threadIdx.x; it < NK * TY * BX; it += C::THREADS) {
      const int yt = it % TY, rest = it / TY, s = rest % NK, exl = rest / NK;
      if (ex0 + exl >= a.nx || y0 + yt >= gy - 1) continue; // no owned output
      const T *ra = sA + (s * TY + yt) * RPA + exl * p;
      const T *rb = sB + (s * TY + yt) * RPA + exl * p;
      T *rc = sC + (s * TY + yt) * RPC + exl * p;
      T *rd = sD + (s * TY + yt) * RPC + exl * p;
      // c = M a;  d = M b + c_x K a, one input line at a time (registers)
      T ua[2 * N - 1], cm[N - 1], ck[N - 1];
#pragma unroll
      for (int j = 0; j < 2 * N - 1; ++j) ua[j] = ra[j];
      line_one<N, T, 0>(a.tab, ua, cm);
#pragma unroll
      for (int i = 0; i < p; ++i) rc[i] = cm[i];
      line_one<N, T, 1>(a.tab, ua, ck);
      T ub[2 * N - 1], dm[N - 1];
#pragma unroll
      for (int j = 0; j < 2 * N - 1; ++j) ub[j] = rb[j];
      line_one<N, T, 0>(a.tab, ub, dm);
#pragma unroll
      for (int i = 0; i < p; ++i) rd[i] = dm[i] + ck[i];
    }
  };

  // phase Z state of this thread's node column
  const int zt = threadIdx.x;
  const int Xt = zt % TX, Yt = zt / TX;
  const int X = x0 + Xt, Y = y0 + Yt;
  // 0: compute, 1: Dirichlet column of the grid's last x or y node (zeros),
  // 2: outside the grid or no column
  const int zone = zt >= C::NZI || X >= gx || Y >= gy ? 2 : (X == gx - 1 || Y == gy - 1 ? 1 : 0);
  // the tile ends exactly at the last (Dirichlet) node: its zeros are
  // written by the threads of the nodes before it
  const bool xtra_x = zone != 2 && Xt == TX - 1 && X + 1 == gx - 1;
  const bool xtra_y = zone != 2 && Yt == TY - 1 && Y + 1 == gy - 1;
  // byte offset of this thread's node column inside a plane (planes hold
  // less than 2^31 bytes: checked on the host)
  const unsigned coff = zone == 2 ? 0u : (unsigned)((Y * gx + X) * (int)sizeof(T));
  T cp = T(0), dp = T(0), yp = T(0);

  // output of node (column offset off) in plane Z (Dirichlet nodes only for
  // the three-term form: the inverse diagonal is 0 there)
  auto put = [&](int Z, unsigned off, T v) {
    const long long zo = (long long)Z * plane; // uniform
    if constexpr (TT) {
      // (a first iterate formed from rhs is 0 on these nodes)
      const long long g = zo + off / (unsigned)sizeof(T);
      const T us = RSRC ? T(0) : src[g];
      const T pv = !RSRC && !RPREV && a.prev ? a.prev[g] : T(0); // null: the zero vector
      dst[g] = us - a.c_tm * (pv - us);
    } else if constexpr (CHEBY == 1) {
      // r = A u - rhs; tm = c_tm tm + c_np diag r; dst = u - tm (five
      // streams: lane addresses, not five uniform bases per plane, which
      // would not fit the scalar registers)
      const long long g = zo + off / (unsigned)sizeof(T);
      const T res = v - a.rhs[g];
      const T t = (a.c_tm != T(0) ? a.c_tm * a.tm[g] : T(0)) + a.c_np * a.diag[g] * res;
      a.tm[g] = t;
      dst[g] = src[g] - t;
    } else {
      __builtin_nontemporal_store(v, &at(dst + zo, off));
    }
  };
  // fused Chebyshev: rhs, diag, tmp (or the previous iterate) and the input
  // at this thread's node column, plane k of a layer in q_*[k].  A rolling
  // prefetch one layer ahead: the first layer's planes are loaded before the
  // layer loop, and as soon as phase Z has written plane k of layer ez, the
  // registers of plane k take the loads of layer ez + 1, in flight during the
  // rest of phase Z and the next layer's phases Y and X (loaded at the start
  // of their own layer, their latency was only partly hidden)
  constexpr int NCH = CHEBY ? p : 1; // CHEBY 5: q_s only
  T q_r[NCH], q_d[CHEBY == 1 ? NCH : 1], q_t[NCH], q_s[NCH];
  double dot_acc = 0.; // CHEBY 5
  auto load_plane = [&](int ez, int k) {
    if constexpr (CHEBY == 1) {
      const long long g = (long long)(ez * p + k) * plane + coff / (unsigned)sizeof(T);
      q_r[k] = a.rhs[g];
      q_d[k] = a.diag[g];
      q_t[k] = a.c_tm != T(0) ? a.tm[g] : T(0);
      q_s[k] = src[g];
    } else if constexpr (DOT) {
      q_s[k] = at(src + (long long)(ez * p + k) * plane, coff);
    } else if constexpr (TT) {
      // uniform plane bases + this lane's 32-bit column offset
      const long long zo = (long long)(ez * p + k) * plane;
      q_r[k] = at(a.rhs + zo, coff);
      if constexpr (!RSRC && !RPREV) q_t[k] = a.prev ? at(a.prev + zo, coff) : T(0);
      if constexpr (!RSRC) q_s[k] = at(src + zo, coff);
    }
  };
  if constexpr (CHEBY != 0) {
    if (zone != 2 && ez_begin < ez_end)
#pragma unroll
      for (int k = 0; k < p; ++k) load_plane(ez_begin, k);
  }
  // three-term form: the inverse diagonal of this node column in plane k of a
  // cell, D = d1 dm[k] + d2 dk[2][k] with the x/y factors d1, d2 of the
  // column (0 off the interior columns); the same in every layer, so formed
  // once
  T dinvk[TT ? p : 1];
  if constexpr (TT) {
    const auto &tb = tptab<N, T>(a.tab);
    const bool inner = zone == 0 && X > 0 && Y > 0;
    const int ix = inner ? X % p : 0, iy = inner ? Y % p : 0;
    const T d1 = tb.dk[0][ix] * tb.dm[iy] + tb.dm[ix] * tb.dk[1][iy];
    const T d2 = tb.dm[ix] * tb.dm[iy];
#pragma unroll
    for (int k = 0; k < p; ++k) dinvk[k] = inner ? T(1) / (d1 * tb.dm[k] + d2 * tb.dk[2][k]) : T(0);
  }
  // output v of this thread's node in plane ez p + k (prefetched streams)
  auto put_own = [&](int ez, int k, T v) {
    if constexpr (TT) {
      const long long zo = (long long)(ez * p + k) * plane; // uniform
      const T dinv = ez * p + k > 0 ? dinvk[k] : T(0);
      const T z1 = a.s_rhs * dinv * q_r[k]; // the first iterate, when formed here
      const T us = RSRC ? z1 : q_s[k];
      const T pv = RPREV ? z1 : (RSRC ? T(0) : q_t[k]);
      const T t = a.c_tm * (pv - us) + a.c_np * dinv * (v - q_r[k]);
      at(dst + zo, coff) = us - t;
      if (ez + 1 < ez_end) load_plane(ez + 1, k);
    } else if constexpr (CHEBY == 1) {
      const long long g = (long long)(ez * p + k) * plane + coff / (unsigned)sizeof(T);
      const T res = v - q_r[k];
      const T t = (a.c_tm != T(0) ? a.c_tm * q_t[k] : T(0)) + a.c_np * q_d[k] * res;
      a.tm[g] = t;
      dst[g] = q_s[k] - t;
      if (ez + 1 < ez_end) load_plane(ez + 1, k);
    } else {
      put(ez * p + k, coff, v);
      if constexpr (DOT) {
        dot_acc += double(q_s[k]) * double(v);
        if (ez + 1 < ez_end) load_plane(ez + 1, k);
      }
    }
  };
  // zero (Dirichlet) outputs of plane Z for this thread: its own node when
  // zone 1 (plane ez p + k), the extra last-node columns
  auto put_zeros = [&](int Z, int ez, int k) {
    if (zone == 1) {
      if (k >= 0) put_own(ez, k, T(0));
      else put(Z, coff, T(0));
    }
    if (xtra_x) put(Z, coff + (unsigned)sizeof(T), T(0));
    if (xtra_y) put(Z, coff + (unsigned)(gx * (int)sizeof(T)), T(0));
    if (xtra_x && xtra_y) put(Z, coff + (unsigned)((gx + 1) * (int)sizeof(T)), T(0));
  };

  // delayed outputs: po[k] of plane ez p + k of the last layer phase Z stored
  T po[DELAY ? N - 1 : 1];
  auto phase_z = [&](int ez, bool store) {
    if (zone == 0) {
      T c[N], d[N], o[N];
      c[0] = cp;
      d[0] = dp;
#pragma unroll
      for (int k = 1; k < N; ++k) {
        c[k] = sC[((k - 1) * TY + Yt) * RPC + Xt];
        d[k] = sD[((k - 1) * TY + Yt) * RPC + Xt];
      }
      T om[N], ok[N];
      eo_apply<N, 1>(tptab<N, T>(a.tab).M, d, om);
#if SFK_TP_SCHED
      __builtin_amdgcn_sched_barrier(0);
#endif
      eo_apply<N, 1>(tptab<N, T>(a.tab).K[2], c, ok);
#pragma unroll
      for (int k = 0; k < N; ++k) o[k] = om[k] + ok[k];
      o[0] += yp;
      if (store) {
        const bool bxy = X == 0 || Y == 0;
#pragma unroll
        for (int k = 0; k < p; ++k) {
          const int Z = ez * p + k;
          if constexpr (DELAY) po[k] = (bxy || Z == 0) ? T(0) : o[k];
          else put_own(ez, k, (bxy || Z == 0) ? T(0) : o[k]);
        }
      }
      cp = c[p];
      dp = d[p];
      yp = o[p];
    }
    if (!DELAY && store && zone != 2) {
#pragma unroll
      for (int k = 0; k < p; ++k) put_zeros(ez * p + k, ez, k);
    }
  };
  // the stores (and, CHEBY 5, the prefetch of the next layer's q_s) of the
  // pending layer.  Vector-memory loads and stores share one completion
  // counter, and these accesses sit in divergent branches, so the wait for
  // the next layer's y lines (store_y) is a wait for everything issued
  // before it: issued at the end of phase Z, the q_s prefetch of CHEBY 5 was
  // waited for at the start of the next layer.  Issued right after that wait
  // instead, it has a whole layer to arrive (Q6 64^3 FP64 CG-Chebyshev: the
  // solve 2 % faster).  The plain matvec and the Chebyshev forms measured
  // slower this way (the held outputs cost registers: Q6 FP64 matvec 337 ->
  // 347 us, Chebyshev steps 583 -> 605 us).
  // (ez: the layer whose outputs are pending, uniform)
  auto flush = [&](int ez) {
    if constexpr (DELAY) {
      if (zone == 0)
#pragma unroll
        for (int k = 0; k < p; ++k) put_own(ez, k, po[k]);
      if (zone != 2)
#pragma unroll
        for (int k = 0; k < p; ++k) put_zeros(ez * p + k, ez, k);
    }
  };

  using one = std::integral_constant<int, 1>;
  using all = std::integral_constant<int, p>;
  // one layer: phase Y from the prefetched registers, the next layer's
  // loads, phases X and Z
  auto layer = [&](int ez, bool store, int ez_next) {
    store_y();
    // the layer before this one stored its outputs (delayed form)
    if (DELAY && store && ez > ez_begin) flush(ez - 1);
    __syncthreads();
    if (ez_next >= 0) load_y(ez_next, 1, all());
    phase_x(all());
    __syncthreads();
    phase_z(ez, store);
  };
  if (ez_begin > 0) {
    // carried state at plane ez_begin p: c, d of the plane below the cell
    // below, then that cell's layer without stores.  The plane's and the
    // layer's y lines are loaded together (two sets of registers): one
    // memory round trip instead of two before the first sweep
    {
      T u0[2 * N - 1];
      bool yact0 = false;
      load_y_to(u0, yact0, ez_begin - 1, 0, one());
      load_y(ez_begin - 1, 1, all());
      store_y_from(u0, yact0);
    }
    __syncthreads();
    phase_x(one());
    __syncthreads();
    if (zone == 0) {
      cp = sC[Yt * RPC + Xt];
      dp = sD[Yt * RPC + Xt];
    }
    layer(ez_begin - 1, false, ez_begin < ez_end ? ez_begin : -1);
  } else if (ez_begin < ez_end) {
    load_y(ez_begin, 1, all());
  }
  for (int ez = ez_begin; ez < ez_end; ++ez) layer(ez, true, ez + 1 < ez_end ? ez + 1 : -1);
  if (DELAY && ez_end > ez_begin) flush(ez_end - 1);
  // the grid's last z plane (Dirichlet)
  if (ez_end == a.nz && zone != 2) {
    const int Z = gz - 1;
    if (zone == 0) put(Z, coff, T(0));
    put_zeros(Z, 0, -1);
  }
  if constexpr (DOT) {
    // workgroup sum of this thread's src . dst terms: within each wave by
    // lane exchanges, then over the waves through LDS (free after the march)
    double v = dot_acc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    double *red = reinterpret_cast<double *>(sA);
    const int wave = threadIdx.x / 64, nw = (C::THREADS + 63) / 64;
    if (threadIdx.x % 64 == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.;
      for (int w = 0; w < nw; ++w) t += red[w];
      a.dot_part[(long long)blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
  }
}

// ------------------------------------------------------------------ host

template <int N, typename T>
static TpTab<N, T> make_tptab(const Tables1D &t, const double inv_jac[3])
{
  // element matrices on [0,1] from the basis values S[i][q] and the basis
  // derivatives G = S D at the Gauss points
  std::vector<double> M(N * N, 0.), K(N * N, 0.), G(N * N, 0.);
  for (int i = 0; i < N; ++i)
    for (int q = 0; q < N; ++q) {
      double g = 0.;
      for (int m = 0; m < N; ++m) g += t.S[i * N + m] * t.D[m * N + q];
      G[i * N + q] = g;
    }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      double m = 0., k = 0.;
      for (int q = 0; q < N; ++q) {
        m += t.weights[q] * t.S[i * N + q] * t.S[j * N + q];
        k += t.weights[q] * G[i * N + q] * G[j * N + q];
      }
      M[i * N + j] = m;
      K[i * N + j] = k;
    }
  // symmetrize (exact in exact arithmetic) and impose the centro-symmetry
  // the even-odd form relies on
  auto sym = [&](std::vector<double> &A) {
    std::vector<double> B(N * N);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j)
        B[i * N + j] = 0.25 * (A[i * N + j] + A[j * N + i] + A[(N - 1 - i) * N + (N - 1 - j)] +
                               A[(N - 1 - j) * N + (N - 1 - i)]);
    A = B;
  };
  sym(M);
  sym(K);
  const double jxw = 1. / (inv_jac[0] * inv_jac[1] * inv_jac[2]);
  TpTab<N, T> tb{};
  tb.M = make_eo<N, T>(M, true);
  for (int d = 0; d < 3; ++d) {
    const double c = inv_jac[d] * inv_jac[d] * jxw;
    std::vector<double> Kd(N * N);
    for (int i = 0; i < N * N; ++i) Kd[i] = c * K[i];
    tb.K[d] = make_eo<N, T>(Kd, true);
    for (int j = 0; j < N; ++j) tb.krow[d][j] = T(Kd[(N - 1) * N + j]);
  }
  for (int j = 0; j < N; ++j) tb.mrow[j] = T(M[(N - 1) * N + j]);
  for (int i = 0; i < N; ++i) {
    // node i of a cell; i = 0 is also node p of the cell before it
    const double m = i == 0 ? M[0] + M[N * N - 1] : M[i * N + i];
    const double k = i == 0 ? K[0] + K[N * N - 1] : K[i * N + i];
    tb.dm[i] = T(m);
    for (int d = 0; d < 3; ++d) tb.dk[d][i] = T(inv_jac[d] * inv_jac[d] * jxw * k);
  }
  return tb;
}

template <int N, typename T>
static hipError_t upload_tp(const Tables1D &t, const double inv_jac[3], void **dev)
{
  const TpTab<N, T> h = make_tptab<N, T>(t, inv_jac);
  hipError_t e = hipMalloc(dev, sizeof(h));
  if (e != hipSuccess) {
    *dev = nullptr;
    return e;
  }
  e = hipMemcpy(*dev, &h, sizeof(h), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(*dev);
    *dev = nullptr;
  }
  return e;
}

hipError_t upload_tp_tables(const Tables1D &t, const double inv_jac[3], bool fp32, void **dev)
{
  switch (t.n) {
#define SFK_CASE(NN) \
  case NN: return fp32 ? upload_tp<NN, float>(t, inv_jac, dev) : upload_tp<NN, double>(t, inv_jac, dev);
    SFK_CASE(2) SFK_CASE(3) SFK_CASE(4) SFK_CASE(5) SFK_CASE(6) SFK_CASE(7) SFK_CASE(8) SFK_CASE(9)
#undef SFK_CASE
  default: return hipErrorInvalidValue;
  }
}

template <int N, typename T, int BX, int BY>
static hipError_t launch_tp_shape(const TpLaunch &L, hipStream_t s)
{
  using C = TpCfg<N, T, BX, BY>;
  constexpr int p = N - 1;
  TpArgs<T> a{};
  a.tab = L.tptab;
  a.src = (const T *)L.src;
  a.dst = (T *)L.dst;
  a.rhs = (const T *)L.rhs;
  a.diag = (const T *)L.diag;
  a.tm = (T *)L.tm;
  a.prev = (const T *)L.prev;
  const bool rsrc = L.cheby == 2 && L.src_from_rhs, rprev = L.cheby == 2 && L.prev_from_rhs;
  if (rsrc && rprev) return hipErrorInvalidValue;
  if (rsrc) a.src = (const T *)L.rhs;
  a.dot_part = L.dot_part;
  a.s_rhs = T(L.s_rhs);
  a.c_np = T(L.c_np);
  a.c_tm = T(L.c_tm);
  a.nx = L.nc[0];
  a.ny = L.nc[1];
  a.nz = L.nc[2];
  a.gx = a.nx * p + 1;
  a.gy = a.ny * p + 1;
  a.gz = a.nz * p + 1;
  a.ntx = (a.nx + BX - 1) / BX;
  a.nty = (a.ny + BY - 1) / BY;
  a.zc = L.zc > 0 ? (L.zc < a.nz ? L.zc : a.nz) : a.nz;
  a.order = L.order;
  a.vstride = L.vstride;
  // 32-bit byte offsets: a plane, and the p+1 planes of a layer, hold less
  // than 2^31 bytes
  if ((long long)N * a.gx * a.gy * (long long)sizeof(T) >= (1LL << 31)) return hipErrorInvalidValue;
  const long long nchunks = (a.nz + a.zc - 1) / a.zc;
  const long long blocks = (long long)a.ntx * a.nty * nchunks;
  if (blocks <= 0 || blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
  const size_t lds = size_t(C::LDS_VALUES) * sizeof(T);
  // the matvec with src . dst (CHEBY 5) exists for FP64 only (the solver's
  // outer matvec)
  decltype(&cont_tp_kernel<N, T, BX, BY, 0>) kdot = nullptr;
  if constexpr (std::is_same<T, double>::value) kdot = cont_tp_kernel<N, T, BX, BY, 5>;
  if (L.dot_part && (L.cheby != 0 || L.nrhs != 1 || !kdot)) return hipErrorInvalidValue;
  auto kern = rsrc ? cont_tp_kernel<N, T, BX, BY, 3>
              : rprev ? cont_tp_kernel<N, T, BX, BY, 4>
              : L.dot_part ? kdot
              : L.cheby == 2 ? cont_tp_kernel<N, T, BX, BY, 2>
                            : (L.cheby == 1 ? cont_tp_kernel<N, T, BX, BY, 1> : cont_tp_kernel<N, T, BX, BY, 0>);
  hipError_t e = allow_dynamic_lds((const void *)kern, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks, L.nrhs), dim3(C::THREADS), lds, s, a);
  return hipGetLastError();
}

// tile shapes (BX, BY) per points per direction N; shape 0 is the default
#define SFK_TP_SHAPES(X)                                                                       \
  X(2, 0, 32, 16) X(2, 1, 64, 8) X(2, 2, 16, 32) X(2, 3, 32, 8) X(2, 4, 128, 4)                 \
  X(3, 0, 16, 8) X(3, 1, 32, 4) X(3, 2, 8, 16) X(3, 3, 16, 4) X(3, 4, 64, 2)                    \
  X(4, 0, 8, 8) X(4, 1, 16, 4) X(4, 2, 4, 16) X(4, 3, 8, 4) X(4, 4, 32, 2)                      \
  X(5, 0, 8, 4) X(5, 1, 16, 2) X(5, 2, 8, 2) X(5, 3, 16, 1) X(5, 4, 32, 1) X(5, 5, 43, 1)                \
  X(6, 0, 4, 4) X(6, 1, 8, 2) X(6, 2, 8, 4) X(6, 3, 4, 2) X(6, 4, 16, 1)                        \
  X(7, 0, 4, 4) X(7, 1, 4, 2) X(7, 2, 8, 2) X(7, 3, 8, 1) X(7, 4, 16, 1) X(7, 5, 24, 1) X(7, 6, 22, 1) \
  X(8, 0, 4, 2) X(8, 1, 2, 4) X(8, 2, 2, 2) X(8, 3, 4, 1) X(8, 4, 8, 1)                         \
  X(9, 0, 4, 2) X(9, 1, 2, 2) X(9, 2, 4, 1) X(9, 3, 2, 1) X(9, 4, 8, 1)

int tp_shape_count(int n)
{
  int c = 0;
#define SFK_X(NN, SH, BXX, BYY) c += n == NN ? 1 : 0;
  SFK_TP_SHAPES(SFK_X)
#undef SFK_X
  return c;
}

void tp_shape(int n, int shape, int *bx, int *by)
{
  *bx = *by = 0;
#define SFK_X(NN, SH, BXX, BYY) \
  if (n == NN && shape == SH) { \
    *bx = BXX;                  \
    *by = BYY;                  \
  }
  SFK_TP_SHAPES(SFK_X)
#undef SFK_X
}

size_t tp_lds_bytes(int n, int shape, bool fp32)
{
  size_t v = 0;
#define SFK_X(NN, SH, BXX, BYY) \
  if (n == NN && shape == SH) v = (size_t)TpCfg<NN, double, BXX, BYY>::LDS_VALUES;
  SFK_TP_SHAPES(SFK_X)
#undef SFK_X
  return v * (fp32 ? 4 : 8);
}

template <typename T>
static hipError_t launch_tp_t(const TpLaunch &L, hipStream_t s)
{
  const int shape = L.shape >= 0 && L.shape < tp_shape_count(L.n) ? L.shape : 0;
#define SFK_X(NN, SH, BXX, BYY) \
  if (L.n == NN && shape == SH) return launch_tp_shape<NN, T, BXX, BYY>(L, s);
  SFK_TP_SHAPES(SFK_X)
#undef SFK_X
  return hipErrorInvalidValue;
}

long long tp_launch_blocks(const TpLaunch &L)
{
  int bx = 0, by = 0;
  tp_shape(L.n, L.shape, &bx, &by);
  if (bx <= 0 || by <= 0) return 0;
  const int nz = L.nc[2], zc = L.zc > 0 ? (L.zc < nz ? L.zc : nz) : nz;
  return (long long)((L.nc[0] + bx - 1) / bx) * ((L.nc[1] + by - 1) / by) * ((nz + zc - 1) / zc);
}

hipError_t launch_cont_tp(const TpLaunch &L, hipStream_t s)
{
  return L.fp32 ? launch_tp_t<float>(L, s) : launch_tp_t<double>(L, s);
}

} // namespace sfk
