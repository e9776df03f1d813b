// Cell-batched sum-factorization kernels for the Laplacian family:
//   Mode::Local  element-local Laplacian (EvaluationCellLaplacian,
//                legacy/include/evaluation_cell_laplacian.h:264-503),
//   Mode::Cont   continuous Q_p Laplacian on a structured node grid with
//                homogeneous Dirichlet rows (colored stores or atomic scatter-add),
//   Mode::DG     DG-SIPG on a Cartesian periodic mesh with the fused
//                Chebyshev update (EvaluationDGLaplacian::do_inner_loop,
//                legacy/include/evaluation_dg_laplacian.h:129-510).
//
// One workgroup processes CB cells; each cell gets N*N threads.  Thread l of
// a cell owns line l of every direction (x-line (j,k), y-line (i,k), z-line
// (i,j), all with index l = hi*N + lo), so each 1D sweep is a register-
// resident even-odd contraction fed by one LDS read and one LDS write per
// point.  Phase order (barriers between phases):
//   L  global -> z-line registers, z value sweep -> LDS A;  DG: neighbour
//      face layers (value + Hermite normal derivative) -> LDS F
//   V1 x value sweep on A;  DG: face arrays, first in-plane sweep
//   V2 y value sweep on A;  DG: face arrays, second in-plane sweep
//   Gx x gradient, cell term, x-face terms, transposed gradient -> B
//   Gy same for y, B += ...
//   Gz same for z, + B, transposed z value sweep -> A
//   By transposed y value sweep
//   Bx transposed x value sweep -> global (DG: fused Chebyshev update;
//      Cont: colored stores or atomic scatter-add)
#include "sfk_device.h"
#include "sfk_kernels.h"

namespace sfk {

enum class Mode { Local = 0, Cont = 1, DG = 2 };

template <int N>
constexpr int cells_per_block()
{
  return 256 / (N * N) > 0 ? 256 / (N * N) : 1;
}
template <int N>
constexpr int threads_per_block()
{
  return ((cells_per_block<N>() * N * N + 63) / 64) * 64;
}
constexpr int kCellMaxThreads = 256;
template <int N, typename T, bool FACES>
constexpr int cell_waves()
{
  constexpr int cb = cells_per_block<N>();
  return lds_waves_per_simd(cb * (2 * N * N * N + (FACES ? 12 * N * N : 0)) * (int)sizeof(T), cb * N * N);
}

template <int N, typename T>
struct CellArgs {
  const DevTab<N, T> *dt; // tables in device memory (read through ctab)
  const T *src;
  T *dst;
  const T *rhs, *diag; // Chebyshev
  T *tm;
  T c_np, c_tm;
  const T *coef;   // per cell q-point coefficient (optional)
  long long ncells;
  int nx, ny, nz;  // cells per direction
  int gx, gy, gz;  // node grid (continuous)
  T cdiag[3];      // ij_d^2 * JxW
  T ij[3];         // 1/h_d
  T jxwf[3];       // JxW of a face normal to d (without weights)
  T tau[3];        // penalty (p+1)^2 / h_d
  int nrhs;
  long long vstride; // elements between right-hand sides
  // continuous: colour of this launch and the row/segment decomposition
  // (a workgroup = CB consecutive cells of one x row; colour = parity of
  // (row y, row z, segment) so that a launch writes every node at most once)
  int color, nseg, nss, nry, nrz;
  int cb;    // cells per workgroup (continuous colored: cells per segment)
  int order; // block order (logical_block)
};

// face terms of SIPG along one line (Cartesian): the thread owns line l of
// direction DIR, q[] are the values at the Gauss points of the line; adds
// value and normal-derivative test contributions to r[] (the transposed
// gradient sweep output).  Signs as in evaluation_dg_laplacian.h:237-252.
template <int N, int DIR, typename T>
__device__ __forceinline__ void sipg_faces(const CellArgs<N, T> &a, const T (&q)[N], const T *F, int l,
                                           T (&r)[N])
{
  constexpr int N2 = N * N;
  const auto &tb = ctab(a.dt);
  T v0 = 0, v1 = 0, d0 = 0, d1 = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    v0 += tb.fv0[i] * q[i];
    v1 += tb.fv1[i] * q[i];
    d0 += tb.fd0[i] * q[i];
    d1 += tb.fd1[i] * q[i];
  }
  const T up0 = F[(2 * DIR) * N2 + l], dp0 = F[(6 + 2 * DIR) * N2 + l];
  const T up1 = F[(2 * DIR + 1) * N2 + l], dp1 = F[(6 + 2 * DIR + 1) * N2 + l];
  const T wf = tb.w[l % N] * tb.w[l / N];
  const T ij = a.ij[DIR], jf = a.jxwf[DIR] * wf, tau = a.tau[DIR];
  const T jump0 = v0 - up0, jump1 = v1 - up1;
  const T tv0 = (T(0.5) * ij * (d0 + dp0) + jump0 * tau) * jf;
  const T td0 = T(0.5) * ij * jf * jump0;
  const T tv1 = (T(-0.5) * ij * (d1 + dp1) + jump1 * tau) * jf;
  const T td1 = T(-0.5) * ij * jf * jump1;
#pragma unroll
  for (int i = 0; i < N; ++i)
    r[i] += tb.fv0[i] * tv0 + tb.fd0[i] * td0 + tb.fv1[i] * tv1 + tb.fd1[i] * td1;
}

// gradient phase of one direction: read Q line, collocation derivative,
// cell term, faces, transposed derivative
template <int N, int DIR, typename T, Mode MODE, bool COEF>
__device__ __forceinline__ void grad_phase(const CellArgs<N, T> &a, const T *Acell, const T *F, int l,
                                           long long cell, T (&r)[N])
{
  const auto &tb = ctab(a.dt);
  T q[N], g[N];
  lds_read_line<N, N, DIR>(Acell, l, q);
  eo_apply<N, -1>(tb.Df, q, g);
  const T c = a.cdiag[DIR];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T f = c * line_weight<N, DIR>(tb.w, l, i);
    if constexpr (COEF) {
      const int dof = line_base<N, N, DIR>(l) + i * line_stride<N, N, DIR>();
      f *= a.coef[cell * (N * N * N) + dof];
    }
    g[i] *= f;
  }
  eo_apply<N, -1>(tb.Db, g, r);
  if constexpr (MODE == Mode::DG) sipg_faces<N, DIR, T>(a, q, F, l, r);
}

// Colored store of a continuous segment.  A workgroup owns CB consecutive
// cells of one x row (a "segment"); the units (row y, row z, segment) are
// coloured by the parity of each index, so two units of one colour share no
// node and a launch writes every node at most once.  The colours are
// launched in ascending order; for every node the unit with the smallest
// colour among those that touch it stores (Dirichlet nodes: 0) and later
// ones add, which makes a zero fill and atomics unnecessary.
//
// The segment's output is (p+1)^2 node rows of ncs*p+1 consecutive nodes
// (ncs cells).  The threads' x lines are merged into those rows in LDS (the
// first x-plane of cell c+1, left in sB by its threads, is added to the last
// x-plane of cell c), and the rows are then written with consecutive lanes
// on consecutive nodes: a store instruction covers a few cache lines
// instead of one row per lane.  Values added to are all loaded before the
// first store.  Called by every thread of the workgroup (barrier inside);
// sA (the A sweep array, dead at this point) holds the rows.
template <int N, typename T>
__device__ __forceinline__ void cont_store_colored(const CellArgs<N, T> &a, T (&o)[N], T *sA, const T *sB, bool act,
                                                   int c, int l, int cy, int cz, int seg, int cx0, int cx_end, T *dst)
{
  constexpr int p = N - 1, N2 = N * N, N3 = N2 * N;
  const int ncs = cx_end - cx0, Lx = ncs * p + 1;
  if (act) {
    if (c + 1 < ncs) o[p] += sB[(c + 1) * N3 + l];
    T *row = sA + l * Lx + c * p;
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (i > 0 || c == 0) row[i] = o[i];
  }
  // per-row terms of the segment's N2 node rows, made once per row instead
  // of once per node: the index of the row's first node, a Dirichlet-row
  // bit (4), and the y/z part of the smallest colour of the units covering
  // the row (rows [ylo,yhi] x [zlo,zhi], each range one or two entries; the
  // segment part is added per node)
  __shared__ long long s_rbase[N2];
  __shared__ int s_rflag[N2];
  if ((int)threadIdx.x < N2) {
    const int rw = (int)threadIdx.x, j = rw % N, k = rw / N;
    const int ylo = (j == 0 && cy > 0) ? cy - 1 : cy, yhi = (j == p && cy + 1 < a.ny) ? cy + 1 : cy;
    const int zlo = (k == 0 && cz > 0) ? cz - 1 : cz, zhi = (k == p && cz + 1 < a.nz) ? cz + 1 : cz;
    const int py = ylo == yhi ? (ylo & 1) : 0, pz = zlo == zhi ? (zlo & 1) : 0;
    const int Y = cy * p + j, Z = cz * p + k;
    const bool bd = Y == 0 || Z == 0 || Y == a.gy - 1 || Z == a.gz - 1;
    s_rbase[rw] = ((long long)Z * a.gy + Y) * a.gx + (long long)cx0 * p;
    s_rflag[rw] = py + 2 * pz + (bd ? 4 : 0);
  }
  __syncthreads();
  // blockDim.x >= CB (p+1)^2 and N2 (ncs p + 1) <= CB N3: N rounds cover the rows
  const int total = N2 * Lx;
  // idx / Lx as (idx * mlx) >> 20 with mlx = ceil(2^20 / Lx): exact while
  // idx * (mlx - 2^20 / Lx) < 2^20 / Lx, i.e. for idx < 2^20 / Lx, which
  // holds here (idx < N blockDim.x <= 2304, Lx <= 8 * 64 + 1)
  const unsigned mlx = ((1u << 20) + unsigned(Lx) - 1u) / unsigned(Lx);
  int mode[N]; // 0 nothing, 1 store, 2 add
  long long gi[N];
  T val[N], old[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const int idx = (int)threadIdx.x + r * (int)blockDim.x;
    mode[r] = 0;
    gi[r] = 0;
    val[r] = T(0);
    old[r] = T(0);
    if (idx < total) {
      const int rw = (int)((unsigned(idx) * mlx) >> 20), xi = idx - rw * Lx;
      const int fl = s_rflag[rw];
      int ps = seg & 1;
      if ((xi == 0 && seg > 0) || (xi == Lx - 1 && seg + 1 < a.nseg)) ps = 0;
      const int m = (fl & 3) + 4 * ps;
      const int X = cx0 * p + xi;
      const bool bd = (fl & 4) || X == 0 || X == a.gx - 1;
      gi[r] = s_rbase[rw] + xi;
      val[r] = bd ? T(0) : sA[idx];
      mode[r] = m == a.color ? 1 : (bd ? 0 : 2);
    }
  }
  // the loads in a loop of their own, so that they are all in flight
  // together (interleaved with the index arithmetic above, the compiler
  // waited for each before the next)
#pragma unroll
  for (int r = 0; r < N; ++r)
    if (mode[r] == 2) old[r] = dst[gi[r]];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if (mode[r] == 1) dst[gi[r]] = val[r];
    else if (mode[r] == 2) dst[gi[r]] = old[r] + val[r];
  }
}

template <int N, typename T, Mode MODE, bool CHEBY, bool REFX, bool COEF>
__global__ void __launch_bounds__(kCellMaxThreads) __attribute__((amdgpu_waves_per_eu(1, cell_waves<N, T, MODE == Mode::DG>())))
cell_kernel(const CellArgs<N, T> a)
{
  constexpr int N2 = N * N, N3 = N * N * N, p = N - 1;
  constexpr bool FACES = MODE == Mode::DG;
  const int CB = a.cb;
  // dynamic LDS: A [CB][N3], B [CB][N3], F [CB][12][N2] (DG only)
  T *sA = dyn_lds<T>();
  T *sB = sA + CB * N3;
  T *sF = sB + CB * N3;

  const int t = threadIdx.x;
  const int c = t / N2, l = t % N2;
  const unsigned blk = logical_block(blockIdx.x, gridDim.x, a.order);
  long long cell0 = (long long)blk * CB;
  // continuous: coordinates of this thread's cell; a colour >= 0 selects the
  // row-segment decomposition, a negative one the cell-linear atomic variant
  int seg = 0, ccx = 0, ccy = 0, ccz = 0, cx0 = 0, cx_end = 0;
  bool act = c < CB && cell0 + c < a.ncells;
  if constexpr (MODE == Mode::Cont) {
    if (a.color >= 0) {
      const int si = (int)(blk % a.nss);
      const int rest = (int)(blk / a.nss);
      seg = (a.color >> 2 & 1) + 2 * si;
      ccy = (a.color & 1) + 2 * (rest % a.nry);
      ccz = (a.color >> 1 & 1) + 2 * (rest / a.nry);
      cx0 = seg * CB;
      cx_end = cx0 + CB < a.nx ? cx0 + CB : a.nx;
      ccx = cx0 + c;
      cell0 = ((long long)ccz * a.ny + ccy) * a.nx + cx0;
      act = c < CB && ccx < cx_end;
    } else {
      // cell0's coordinates are uniform; cell0 + c is c steps further in x
      // unless that crosses the end of the row
      CellXYZ q = traverse_xyz((unsigned)cell0, a.nx, a.ny, a.nz, 0);
      if (q.x + c < (unsigned)a.nx) q.x += c;
      else q = traverse_xyz((unsigned)(cell0 + c), a.nx, a.ny, a.nz, 0);
      ccx = (int)q.x;
      ccy = (int)q.y;
      ccz = (int)q.z;
    }
  }
  // DG: the cell at this traversal position and its periodic face neighbours
  int nbr[6] = {0, 0, 0, 0, 0, 0};
  long long cell = cell0 + c;
  if constexpr (MODE == Mode::DG) {
    if (act) cell = locate_cell((unsigned)cell0, (unsigned)c, a.nx, a.ny, a.nz, a.order, nbr);
  }
  T *A = sA + (act ? c : 0) * N3;
  T *B = sB + (act ? c : 0) * N3;
  T *F = sF + (FACES && act ? c * 12 * N2 : 0);

  // right-hand side rr = blockIdx.y (a loop over the right-hand sides here
  // would keep every 1D table live across it and spill scalar registers)
  {
    const long long rr = blockIdx.y;
    const T *src = a.src + rr * a.vstride;
    T *dst = a.dst + rr * a.vstride;
    // ---------------------------------------------------------------- L
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      if constexpr (MODE == Mode::Cont) {
        // (staging the segment's node rows through LDS for coalesced loads
        // measured slower: the extra barrier costs more than it saves)
        const int X = ccx * p + l % N, Y = ccy * p + l / N;
        const bool bxy = X == 0 || Y == 0 || X == a.gx - 1 || Y == a.gy - 1;
        const long long g0 = (long long)ccz * p * a.gy * a.gx + (long long)Y * a.gx + X;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const int Z = ccz * p + k;
          const bool bd = bxy || Z == 0 || Z == a.gz - 1;
          u[k] = bd ? T(0) : src[g0 + (long long)k * a.gy * a.gx];
        }
      } else {
#pragma unroll
        for (int k = 0; k < N; ++k) u[k] = src[cell * N3 + k * N2 + l];
      }
      eo_apply<N, 1>(tb.Sf, u, v);
      lds_write_line<N, N, 2>(A, l, v);
    }
    if constexpr (FACES) {
      // neighbour layers: value layer and Hermite normal derivative
      // w0*(inner - face) (evaluation_dg_laplacian.h:196-203, 272-283);
      // thread l of a cell loads face point l of all six faces (twelve
      // independent loads, issued together)
      if (act) {
        T uf[6], ui[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          const int d = f / 2, s = f % 2;
          const T *nb = src + (long long)nbr[f] * N3;
          const int lf = s ? 0 : p, li = s ? 1 : p - 1;
          const int of = d == 0 ? face_dof<N, 0>(lf, l) : (d == 1 ? face_dof<N, 1>(lf, l) : face_dof<N, 2>(lf, l));
          const int oi = d == 0 ? face_dof<N, 0>(li, l) : (d == 1 ? face_dof<N, 1>(li, l) : face_dof<N, 2>(li, l));
          uf[f] = nb[of];
          ui[f] = nb[oi];
        }
        const T hd = ctab(a.dt).hd;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          F[f * N2 + l] = uf[f];
          F[(6 + f) * N2 + l] = (f % 2 ? -hd : hd) * (ui[f] - uf[f]);
        }
      }
    }
    __syncthreads();
    // --------------------------------------------------------------- V1
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      lds_read_line<N, N, 0>(A, l, u);
      eo_apply<N, 1>(tb.Sf, u, v);
      lds_write_line<N, N, 0>(A, l, v);
      if constexpr (FACES) {
        for (int it = l; it < 12 * N; it += N2) {
          T *fl = F + (it / N) * N2 + (it % N) * N; // lo-direction line
#pragma unroll
          for (int i = 0; i < N; ++i) u[i] = fl[i];
          eo_apply<N, 1>(tb.Sf, u, v);
#pragma unroll
          for (int i = 0; i < N; ++i) fl[i] = v[i];
        }
      }
    }
    __syncthreads();
    // --------------------------------------------------------------- V2
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      lds_read_line<N, N, 1>(A, l, u);
      eo_apply<N, 1>(tb.Sf, u, v);
      lds_write_line<N, N, 1>(A, l, v);
      if constexpr (FACES) {
        for (int it = l; it < 12 * N; it += N2) {
          T *fl = F + (it / N) * N2 + (it % N); // hi-direction line
#pragma unroll
          for (int i = 0; i < N; ++i) u[i] = fl[i * N];
          eo_apply<N, 1>(tb.Sf, u, v);
#pragma unroll
          for (int i = 0; i < N; ++i) fl[i * N] = v[i];
        }
      }
    }
    __syncthreads();
    // --------------------------------------------------------------- Gx
    if (act) {
      const auto &tb = ctab(a.dt);
      T r[N];
      grad_phase<N, 0, T, MODE, COEF>(a, A, F, l, cell, r);
      lds_write_line<N, N, 0>(B, l, r);
    }
    __syncthreads();
    // --------------------------------------------------------------- Gy
    if (act) {
      const auto &tb = ctab(a.dt);
      T r[N], b[N];
      grad_phase<N, 1, T, MODE, COEF>(a, A, F, l, cell, r);
      lds_read_line<N, N, 1>(B, l, b);
#pragma unroll
      for (int i = 0; i < N; ++i) r[i] += b[i];
      lds_write_line<N, N, 1>(B, l, r);
    }
    __syncthreads();
    // --------------------------------------------------------------- Gz
    if (act) {
      const auto &tb = ctab(a.dt);
      T r[N], b[N], o[N];
      grad_phase<N, 2, T, MODE, COEF>(a, A, F, l, cell, r);
      lds_read_line<N, N, 2>(B, l, b);
#pragma unroll
      for (int i = 0; i < N; ++i) r[i] += b[i];
      eo_apply<N, 1>(tb.Sb, r, o);
      lds_write_line<N, N, 2>(A, l, o);
    }
    __syncthreads();
    // --------------------------------------------------------------- By
    // DG Chebyshev: this thread's x line of rhs, diag, tmp and the input,
    // loaded one phase ahead of the update that uses them
    T crhs[CHEBY ? N : 1], cdiag[CHEBY ? N : 1], ctm[CHEBY ? N : 1], csrc[CHEBY ? N : 1];
    if (act) {
      if constexpr (MODE == Mode::DG && CHEBY) {
        const long long g0 = cell * N3 + l * N;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          crhs[i] = a.rhs[g0 + i];
          cdiag[i] = a.diag[g0 + i];
          ctm[i] = a.c_tm != T(0) ? a.tm[g0 + i] : T(0); // c_tm == 0: tm not read
          csrc[i] = src[g0 + i];
        }
      }
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      lds_read_line<N, N, 1>(A, l, u);
      eo_apply<N, 1>(tb.Sb, u, v);
      lds_write_line<N, N, 1>(A, l, v);
    }
    __syncthreads();
    // --------------------------------------------------------------- Bx
    T o[N];
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N];
      lds_read_line<N, N, 0>(A, l, u);
      if constexpr (REFX)
        eo_apply<N, 1>(tb.Sf, u, o); // evaluation_dg_laplacian.h:481-483
      else
        eo_apply<N, 1>(tb.Sb, u, o);
      if constexpr (MODE == Mode::Cont) {
        // x-merge: the first x-plane of cell c+1 is the last one of cell c
        if (a.color >= 0) B[l] = o[0];
      }
    }
    if constexpr (MODE == Mode::Cont) {
      if (a.color >= 0) {
        __syncthreads();
        cont_store_colored<N, T>(a, o, sA, sB, act, c, l, ccy, ccz, seg, cx0, cx_end, dst);
      } else if (act) {
        const int Y = ccy * p + l % N, Z = ccz * p + l / N;
        const bool byz = Y == 0 || Z == 0 || Y == a.gy - 1 || Z == a.gz - 1;
        const long long g0 = ((long long)Z * a.gy + Y) * a.gx + (long long)ccx * p;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const int X = ccx * p + i;
          if (!(byz || X == 0 || X == a.gx - 1)) atomicAdd(dst + g0 + i, o[i]);
        }
      }
    }
    if (act) {
      const auto &tb = ctab(a.dt);
      if constexpr (MODE == Mode::Cont) {
      } else if constexpr (CHEBY) {
        // r = A u - rhs; tm = c_tm*tm + c_np*diag*r; dst = u - tm
        // (evaluation_dg_laplacian.h:485-502)
        const long long g0 = cell * N3 + l * N;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const long long g = g0 + i;
          const T res = o[i] - crhs[i];
          const T base = a.c_np * cdiag[i] * res;
          const T tmv = a.c_tm != T(0) ? a.c_tm * ctm[i] + base : base;
          a.tm[g] = tmv;
          dst[g] = csrc[i] - tmv;
        }
      } else {
        const long long g0 = cell * N3 + l * N;
        // written once, read by nobody in this launch: non-temporal stores
#pragma unroll
        for (int i = 0; i < N; ++i) __builtin_nontemporal_store(o[i], dst + g0 + i);
      }
    }
  }
}

// Diagonal of the Cartesian Laplacian (local or continuous, optionally with
// a per-q coefficient): diag_i = sum_d sum_q c_d k_q w_q (dphi_i/dxi_d)^2
// prod_{e!=d} phi_i^2 -- a sum-factorized contraction of the weights with
// the squared 1D matrices (G1 = S D: basis derivative at Gauss points).
template <int N, typename T>
struct DiagArgs {
  Eo<N, T> S2b, G2b; // backward contractions with S^2 and G1^2
  Eo<N, T> S1b;      // load vector: backward contraction with S
  int load;          // 1: b_i = sum_q f JxW_q phi_i(x_q) instead of the diagonal
  T fval, jxw;
  T w[N];
  const T *coef;
  T *diag;
  long long ncells;
  int nx, ny, nz, gx, gy, gz;
  T cdiag[3];
  int cont;
};

template <int N, typename T>
__global__ void __launch_bounds__(threads_per_block<N>()) diag_kernel(const DiagArgs<N, T> a)
{
  constexpr int CB = cells_per_block<N>();
  constexpr int N2 = N * N, N3 = N * N * N, p = N - 1;
  __shared__ T sA[CB * N3];
  __shared__ T sR[CB * N3];
  const int t = threadIdx.x, c = t / N2, l = t % N2;
  const long long cell = (long long)blockIdx.x * CB + c;
  const bool act = c < CB && cell < a.ncells;
  T *A = sA + (act ? c : 0) * N3;
  T *R = sR + (act ? c : 0) * N3;
  for (int d = 0; d < (a.load ? 1 : 3); ++d) {
    if (act) {
      // weights c_d w_q k_q on x-line l
      T u[N], v[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T f = (a.load ? a.fval * a.jxw : a.cdiag[d]) * line_weight<N, 0>(a.w, l, i);
        if (a.coef && !a.load) f *= a.coef[cell * N3 + l * N + i];
        u[i] = f;
      }
      if (a.load) eo_apply<N, 1>(a.S1b, u, v);
      else if (d == 0) eo_apply<N, 1>(a.G2b, u, v);
      else eo_apply<N, 1>(a.S2b, u, v);
      lds_write_line<N, N, 0>(A, l, v);
    }
    __syncthreads();
    if (act) {
      T u[N], v[N];
      lds_read_line<N, N, 1>(A, l, u);
      if (a.load) eo_apply<N, 1>(a.S1b, u, v);
      else if (d == 1) eo_apply<N, 1>(a.G2b, u, v);
      else eo_apply<N, 1>(a.S2b, u, v);
      lds_write_line<N, N, 1>(A, l, v);
    }
    __syncthreads();
    if (act) {
      T u[N], v[N], r[N];
      lds_read_line<N, N, 2>(A, l, u);
      if (a.load) eo_apply<N, 1>(a.S1b, u, v);
      else if (d == 2) eo_apply<N, 1>(a.G2b, u, v);
      else eo_apply<N, 1>(a.S2b, u, v);
      if (d > 0) lds_read_line<N, N, 2>(R, l, r);
#pragma unroll
      for (int i = 0; i < N; ++i) r[i] = d > 0 ? r[i] + v[i] : v[i];
      lds_write_line<N, N, 2>(R, l, r);
    }
    __syncthreads();
  }
  if (act) {
    T r[N];
    lds_read_line<N, N, 2>(R, l, r); // z-line l = j*N + i
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (a.cont) {
        const int cx = (int)(cell % a.nx);
        const long long rr = cell / a.nx;
        const int cy = (int)(rr % a.ny), cz = (int)(rr / a.ny);
        const int X = cx * p + l % N, Y = cy * p + l / N, Z = cz * p + k;
        const bool bd = X == 0 || Y == 0 || Z == 0 || X == a.gx - 1 || Y == a.gy - 1 || Z == a.gz - 1;
        if (!bd) atomicAdd(a.diag + ((long long)Z * a.gy + Y) * a.gx + X, r[k]);
      } else {
        a.diag[cell * N3 + k * N2 + l] = r[k];
      }
    }
  }
}

// Gradient of lambda^T A(k) u with respect to the per-q coefficient of the
// continuous Laplacian: g_q = sum_rhs JxW_q sum_d ij_d^2 (du/dxi_d)(dlambda/dxi_d)
// at every cell quadrature point.
template <int N, typename T>
struct CoefGradArgs {
  const DevTab<N, T> *dt;
  const T *u, *lam;
  T *grad;
  long long ncells;
  int nx, ny, nz, gx, gy, gz;
  T cdiag[3];
  int nrhs;
  long long vstride;
};

template <int N, typename T>
__global__ void __launch_bounds__(threads_per_block<N>()) coef_grad_kernel(const CoefGradArgs<N, T> a)
{
  constexpr int CB = cells_per_block<N>();
  constexpr int N2 = N * N, N3 = N * N * N, p = N - 1;
  __shared__ T sU[CB * N3];
  __shared__ T sL[CB * N3];
  __shared__ T sG[CB * N3];
  const int t = threadIdx.x, c = t / N2, l = t % N2;
  const long long cell = (long long)blockIdx.x * CB + c;
  const bool act = c < CB && cell < a.ncells;
  T *U = sU + (act ? c : 0) * N3;
  T *Lm = sL + (act ? c : 0) * N3;
  T *G = sG + (act ? c : 0) * N3;
  const int cx = act ? (int)(cell % a.nx) : 0;
  const long long rr0 = act ? cell / a.nx : 0;
  const int cy = (int)(rr0 % a.ny), cz = (int)(rr0 / a.ny);
  for (int r = 0; r < a.nrhs; ++r) {
    if (act) {
      T u[N], v[N], w[N], z[N];
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const int X = cx * p + l % N, Y = cy * p + l / N, Z = cz * p + k;
        const bool bd = X == 0 || Y == 0 || Z == 0 || X == a.gx - 1 || Y == a.gy - 1 || Z == a.gz - 1;
        const long long g = ((long long)Z * a.gy + Y) * a.gx + X + r * a.vstride;
        u[k] = bd ? T(0) : a.u[g];
        w[k] = bd ? T(0) : a.lam[g];
      }
      eo_apply<N, 1>(ctab(a.dt).Sf, u, v);
      eo_apply<N, 1>(ctab(a.dt).Sf, w, z);
      lds_write_line<N, N, 2>(U, l, v);
      lds_write_line<N, N, 2>(Lm, l, z);
    }
    __syncthreads();
    if (act) {
      T u[N], v[N];
      lds_read_line<N, N, 0>(U, l, u);
      eo_apply<N, 1>(ctab(a.dt).Sf, u, v);
      lds_write_line<N, N, 0>(U, l, v);
      lds_read_line<N, N, 0>(Lm, l, u);
      eo_apply<N, 1>(ctab(a.dt).Sf, u, v);
      lds_write_line<N, N, 0>(Lm, l, v);
    }
    __syncthreads();
    if (act) {
      T u[N], v[N];
      lds_read_line<N, N, 1>(U, l, u);
      eo_apply<N, 1>(ctab(a.dt).Sf, u, v);
      lds_write_line<N, N, 1>(U, l, v);
      lds_read_line<N, N, 1>(Lm, l, u);
      eo_apply<N, 1>(ctab(a.dt).Sf, u, v);
      lds_write_line<N, N, 1>(Lm, l, v);
    }
    __syncthreads();
    auto dir_term = [&](auto dirc, bool first) {
      constexpr int DIR = decltype(dirc)::value;
      T u[N], w[N], gu[N], gw[N], acc[N];
      lds_read_line<N, N, DIR>(U, l, u);
      lds_read_line<N, N, DIR>(Lm, l, w);
      eo_apply<N, -1>(ctab(a.dt).Df, u, gu);
      eo_apply<N, -1>(ctab(a.dt).Df, w, gw);
      if (first) {
#pragma unroll
        for (int i = 0; i < N; ++i) acc[i] = T(0);
      } else {
        lds_read_line<N, N, DIR>(G, l, acc);
      }
#pragma unroll
      for (int i = 0; i < N; ++i) acc[i] += a.cdiag[DIR] * line_weight<N, DIR>(ctab(a.dt).w, l, i) * gu[i] * gw[i];
      lds_write_line<N, N, DIR>(G, l, acc);
    };
    if (act) dir_term(std::integral_constant<int, 0>(), r == 0);
    __syncthreads();
    if (act) dir_term(std::integral_constant<int, 1>(), false);
    __syncthreads();
    if (act) dir_term(std::integral_constant<int, 2>(), false);
    __syncthreads();
  }
  if (act) {
    T v[N];
    lds_read_line<N, N, 0>(G, l, v);
#pragma unroll
    for (int i = 0; i < N; ++i) a.grad[cell * N3 + l * N + i] = v[i];
  }
}


template <int N, typename T>
static hipError_t launch_coef_grad_n(const CoefGradLaunch &L, hipStream_t s)
{
  CoefGradArgs<N, T> a{};
  a.dt = (const DevTab<N, T> *)L.dtab;
  a.u = (const T *)L.u;
  a.lam = (const T *)L.lambda;
  a.grad = (T *)L.grad;
  a.ncells = L.ncells;
  a.nx = L.nc[0];
  a.ny = L.nc[1];
  a.nz = L.nc[2];
  a.gx = L.nc[0] * (N - 1) + 1;
  a.gy = L.nc[1] * (N - 1) + 1;
  a.gz = L.nc[2] * (N - 1) + 1;
  const double jxw = 1. / (L.inv_jac[0] * L.inv_jac[1] * L.inv_jac[2]);
  for (int d = 0; d < 3; ++d) a.cdiag[d] = T(L.inv_jac[d] * L.inv_jac[d] * jxw);
  a.nrhs = L.nrhs;
  a.vstride = L.vstride;
  constexpr int CB = cells_per_block<N>();
  hipLaunchKernelGGL((coef_grad_kernel<N, T>), dim3((unsigned)((L.ncells + CB - 1) / CB)),
                     dim3(threads_per_block<N>()), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------- launch

// continuous scatter schedule: colored (no atomics, no zero fill) unless the
// caller asks for the atomic variant or, by default, each of the 8 colour
// launches would hold fewer workgroups than the chip has CUs
static bool cont_colored(const CellLaunch &L, int cb)
{
  if (L.cont_scatter >= 0) return L.cont_scatter == 1;
  return L.ncells / cb >= 8 * 256;
}

template <int N, typename T>
static hipError_t launch_cell_n(const CellLaunch &L, hipStream_t stream)
{
  CellArgs<N, T> a{};
  a.dt = (const DevTab<N, T> *)L.dtab;
  a.src = (const T *)L.src;
  a.dst = (T *)L.dst;
  a.rhs = (const T *)L.rhs;
  a.diag = (const T *)L.diag;
  a.tm = (T *)L.tm;
  a.c_np = T(L.c_np);
  a.c_tm = T(L.c_tm);
  a.coef = (const T *)L.coef;
  a.ncells = L.ncells;
  a.nx = L.nc[0];
  a.ny = L.nc[1];
  a.nz = L.nc[2];
  a.gx = L.nc[0] * (N - 1) + 1;
  a.gy = L.nc[1] * (N - 1) + 1;
  a.gz = L.nc[2] * (N - 1) + 1;
  const double jxw = 1. / (L.inv_jac[0] * L.inv_jac[1] * L.inv_jac[2]);
  for (int d = 0; d < 3; ++d) {
    a.cdiag[d] = T(L.inv_jac[d] * L.inv_jac[d] * jxw);
    a.ij[d] = T(L.inv_jac[d]);
    a.jxwf[d] = T(jxw * L.inv_jac[d]);
    a.tau[d] = T(double(N * N) * L.inv_jac[d]);
  }
  a.nrhs = L.nrhs;
  a.vstride = L.vstride;
  const int CB = L.cb > 0 ? L.cb : cells_per_block<N>();
  if (CB * N * N > kCellMaxThreads) return hipErrorInvalidConfiguration;
  a.cb = CB;
  a.order = L.order;
  constexpr int N3 = N * N * N;
  const size_t lds = size_t(CB) * (2 * N3 + (L.mode == 2 ? 12 * N * N : 0)) * sizeof(T);
  const unsigned blocks = (unsigned)((L.ncells + CB - 1) / CB);
  const dim3 grid(blocks, L.nrhs), block(((CB * N * N + 63) / 64) * 64);
  auto launch = [&](auto kern, dim3 g) {
    const hipError_t e = allow_dynamic_lds((const void *)kern, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, g, block, lds, stream, a);
    return hipGetLastError();
  };
  if (L.mode == 0) {
    return launch(L.coef ? cell_kernel<N, T, Mode::Local, false, false, true>
                         : cell_kernel<N, T, Mode::Local, false, false, false>, grid);
  } else if (L.mode == 1) {
    auto kern = L.coef ? cell_kernel<N, T, Mode::Cont, false, false, true>
                       : cell_kernel<N, T, Mode::Cont, false, false, false>;
    if (!cont_colored(L, CB)) {
      a.color = -1;
      const hipError_t e = vec_fill(std::is_same<T, float>::value, L.dst, L.vstride * L.nrhs, 0., stream);
      if (e != hipSuccess) return e;
      return launch(kern, grid);
    }
    a.nseg = (a.nx + CB - 1) / CB;
    // colours in ascending order: the first unit to touch a node stores it
    for (int color = 0; color < 8; ++color) {
      const int ps = color >> 2 & 1, py = color & 1, pz = color >> 1 & 1;
      a.color = color;
      a.nss = (a.nseg - ps + 1) / 2;
      a.nry = (a.ny - py + 1) / 2;
      a.nrz = (a.nz - pz + 1) / 2;
      const long long nb = (long long)a.nss * a.nry * a.nrz;
      if (nb == 0) continue;
      const hipError_t e = launch(kern, dim3((unsigned)nb, L.nrhs));
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  if (L.cheby)
    return launch(L.refx ? cell_kernel<N, T, Mode::DG, true, true, false> : cell_kernel<N, T, Mode::DG, true, false, false>,
                  grid);
  return launch(L.refx ? cell_kernel<N, T, Mode::DG, false, true, false> : cell_kernel<N, T, Mode::DG, false, false, false>,
                grid);
}

template <int N, typename T>
static hipError_t launch_diag_n(const CellLaunch &L, hipStream_t stream)
{
  DiagArgs<N, T> a{};
  const Tables1D &t = *L.tables;
  std::vector<double> S2(N * N), G2(N * N);
  for (int i = 0; i < N; ++i)
    for (int q = 0; q < N; ++q) {
      double g = 0.;
      for (int m = 0; m < N; ++m) g += t.S[i * N + m] * t.D[m * N + q];
      S2[i * N + q] = t.S[i * N + q] * t.S[i * N + q];
      G2[i * N + q] = g * g;
    }
  a.S2b = make_eo<N, T>(S2, false);
  a.G2b = make_eo<N, T>(G2, false);
  a.S1b = make_eo<N, T>(t.S, false);
  a.load = L.load ? 1 : 0;
  a.fval = T(L.load_value);
  for (int i = 0; i < N; ++i) a.w[i] = T(t.weights[i]);
  a.coef = (const T *)L.coef;
  a.diag = (T *)L.dst;
  a.ncells = L.ncells;
  a.nx = L.nc[0];
  a.ny = L.nc[1];
  a.nz = L.nc[2];
  a.gx = L.nc[0] * (N - 1) + 1;
  a.gy = L.nc[1] * (N - 1) + 1;
  a.gz = L.nc[2] * (N - 1) + 1;
  const double jxw = 1. / (L.inv_jac[0] * L.inv_jac[1] * L.inv_jac[2]);
  for (int d = 0; d < 3; ++d) a.cdiag[d] = T(L.inv_jac[d] * L.inv_jac[d] * jxw);
  a.jxw = T(jxw);
  a.cont = L.mode == 1;
  constexpr int CB = cells_per_block<N>();
  const unsigned blocks = (unsigned)((L.ncells + CB - 1) / CB);
  hipLaunchKernelGGL((diag_kernel<N, T>), dim3(blocks), dim3(threads_per_block<N>()), 0, stream, a);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatch_cell(const CellLaunch &L, hipStream_t s, bool diag)
{
  switch (L.tables->n) {
#define SFK_CASE(NN) \
  case NN: return diag ? launch_diag_n<NN, T>(L, s) : launch_cell_n<NN, T>(L, s);
    SFK_CASE(2) SFK_CASE(3) SFK_CASE(4) SFK_CASE(5) SFK_CASE(6) SFK_CASE(7) SFK_CASE(8) SFK_CASE(9) \
    SFK_CASE(10) SFK_CASE(11) SFK_CASE(12) SFK_CASE(13) SFK_CASE(14) SFK_CASE(15) SFK_CASE(16)
#undef SFK_CASE
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_cell(const CellLaunch &L, hipStream_t stream)
{
  return L.fp32 ? dispatch_cell<float>(L, stream, false) : dispatch_cell<double>(L, stream, false);
}

hipError_t launch_cell_diagonal(const CellLaunch &L, hipStream_t stream)
{
  return L.fp32 ? dispatch_cell<float>(L, stream, true) : dispatch_cell<double>(L, stream, true);
}

template <typename T>
static hipError_t dispatch_coef_grad(const CoefGradLaunch &L, hipStream_t s)
{
  switch (L.tables->n) {
#define SFK_CASE(NN) \
  case NN: return launch_coef_grad_n<NN, T>(L, s);
    SFK_CASE(2) SFK_CASE(3) SFK_CASE(4) SFK_CASE(5) SFK_CASE(6) SFK_CASE(7) SFK_CASE(8) SFK_CASE(9) \
    SFK_CASE(10) SFK_CASE(11) SFK_CASE(12) SFK_CASE(13) SFK_CASE(14) SFK_CASE(15) SFK_CASE(16)
#undef SFK_CASE
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_coef_gradient(const CoefGradLaunch &L, hipStream_t stream)
{
  return L.fp32 ? dispatch_coef_grad<float>(L, stream) : dispatch_coef_grad<double>(L, stream);
}

} // namespace sfk
