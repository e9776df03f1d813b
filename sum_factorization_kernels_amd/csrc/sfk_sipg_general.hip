// DG-SIPG with a per-quadrature-point metric (deformed periodic mesh) and
// the device-side generator of the metric of a sine-deformed mesh.
//
// Operator (the reference has the Cartesian special case only,
// legacy/include/evaluation_dg_laplacian.h:129-510):
//   a(u,v) = sum_K (G grad^u, grad^v)_K
//          + sum_F [ -{du/dn}[v] - {dv/dn}[u] + sigma [u][v] ]_F
// with G = JxW J^{-1}J^{-T} per cell q-point and, per face q-point,
// aX = JxW_f J_X^{-1} n for both sides and sigma = tau JxW_f (layouts in
// include/sfk.h).  The neighbour's value and normal derivative come from
// its two nodal layers next to the face (Hermite-like basis,
// evaluation_dg_laplacian.h:196-203), its tangential derivatives from the
// in-plane collocation derivative of the interpolated face values.  All face
// test functions, value and gradient, are injected at the cell's Gauss
// points through l_q(s), so one transposed collocation-derivative sweep per
// direction integrates cell and face terms together.
//
// One workgroup = CB cells, N*N threads per cell, thread l owns line l of
// every direction.  Phases (barrier after each):
//   L   u z-lines -> values; neighbour layers -> F (value, normal derivative)
//   V1  x value sweep; F first in-plane sweep
//   V2  y value sweep; F second in-plane sweep
//   G   reference gradient Gx, Gy, Gz (one line of each per thread)
//   E   traces of u and grad u on all 6 faces, neighbour data, face metric
//       -> face test coefficients kept in registers (tv, td[3] per face)
//   M   z-line owner: T = G grad^u (cell metric) + z-face gradient tests
//   Iy  y-line owner: y-face gradient tests into T
//   Ix  x-line owner: x-face gradient tests, x transposed sweep + x value
//       tests -> R
//   By  y transposed sweep + y value tests + R -> R
//   Bz  z transposed sweep + z value tests + R, z value sweep^T -> Q
//   Sy, Sx  transposed value sweeps; output (+ fused Chebyshev update)
#include "sfk_device.h"
#include "sfk_kernels.h"

namespace sfk {

template <int N>
constexpr int gen_cells_per_block()
{
  // default (4 N^3 + 12 N^2 values of LDS per cell); N = 6: 3 cells (108
  // threads in two waves) measured fastest on MI355X for the 32^3 workload
  return N <= 3 ? 256 / (N * N) : (N == 4 ? 12 : (N == 5 ? 8 : (N == 6 ? 3 : (N == 7 ? 3 : (N == 8 ? 2 : 1)))));
}
constexpr int kGenMaxThreads = 256;
template <int N, typename T, bool DIAG>
constexpr int gen_waves()
{
  constexpr int cb = gen_cells_per_block<N>();
  return lds_waves_per_simd(cb * (4 * N * N * N + (DIAG ? 0 : 12 * N * N)) * (int)sizeof(T), cb * N * N);
}

template <int N, typename T>
struct GenArgs {
  const DevTab<N, T> *dt; // tables in device memory (read through ctab/cdevtab)
  const T *src;
  T *dst;
  const T *rhs, *diag;
  T *tm;
  T c_np, c_tm;
  const T *cm; // [cell][6][q]
  const T *fm; // [d][cell][7][fa]
  long long ncells;
  long long nitems; // cells (matvec) or cells * N^3 (diagonal)
  int nx, ny, nz;
  int cb;    // cells (items) per workgroup
  int order; // block order (logical_block)
};

// face metric of the two faces of direction DIR: fmv[0] of the lower face
// (stored with the lower neighbour `lower`), fmv[1] of the upper face
template <int N, int DIR, typename T>
__device__ __forceinline__ void load_face_metric(const GenArgs<N, T> &a, int lower, long long cell, int l,
                                                 T (&fmv)[2][7])
{
  constexpr int N2 = N * N;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const T *fmp = a.fm + ((long long)DIR * a.ncells + (s ? cell : (long long)lower)) * 7 * N2 + l;
#pragma unroll
    for (int e = 0; e < 7; ++e) fmv[s][e] = fmp[e * N2];
  }
}

// traces, neighbour data and face metric of one direction -> test coefficients
template <int N, int DIR, typename T, bool DIAG>
__device__ __forceinline__ void face_terms(const GenArgs<N, T> &a, const T *Q, const T *Gx, const T *Gy, const T *Gz,
                                           const T *F, int l, const T (&dlo)[N], const T (&dhi)[N],
                                           const T (&fmv)[2][7], T (&tv)[2], T (&td)[2][3])
{
  constexpr int N2 = N * N;
  const auto &tb = ctab(a.dt);
  T q[N], g[3][N];
  lds_read_line<N, N, DIR>(Q, l, q);
  lds_read_line<N, N, DIR>(Gx, l, g[0]);
  lds_read_line<N, N, DIR>(Gy, l, g[1]);
  lds_read_line<N, N, DIR>(Gz, l, g[2]);
  constexpr int LO = DIR == 0 ? 1 : 0, HI = DIR == 2 ? 1 : 2;
  const int llo = l % N, lhi = l / N;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const auto *fv = s ? tb.fv1 : tb.fv0;
    const auto *fd = s ? tb.fd1 : tb.fd0;
    T um = 0, gm[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < N; ++i) {
      um += fv[i] * q[i];
      gm[DIR] += fd[i] * q[i];
      gm[LO] += fv[i] * g[LO][i];
      gm[HI] += fv[i] * g[HI][i];
    }
    T up = 0, gp[3] = {0, 0, 0};
    if constexpr (!DIAG) {
      const T *FV = F + (2 * DIR + s) * N2;
      up = FV[l];
      gp[DIR] = F[(6 + 2 * DIR + s) * N2 + l];
#pragma unroll
      for (int m = 0; m < N; ++m) {
        gp[LO] += dlo[m] * FV[lhi * N + m];
        gp[HI] += dhi[m] * FV[m * N + llo];
      }
    }
    T ao[3], an[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      ao[e] = s ? fmv[s][e] : fmv[s][3 + e];
      an[e] = s ? fmv[s][3 + e] : fmv[s][e];
    }
    const T sigma = fmv[s][6];
    const T sn = s ? T(1) : T(-1);
    T dn = 0;
#pragma unroll
    for (int e = 0; e < 3; ++e) dn += ao[e] * gm[e] + an[e] * gp[e];
    const T jump = um - up;
    tv[s] = T(-0.5) * sn * dn + sigma * jump;
#pragma unroll
    for (int e = 0; e < 3; ++e) td[s][e] = T(-0.5) * jump * sn * ao[e];
  }
}

// inject the gradient tests of direction DIR's faces into T (own d-line)
template <int N, int DIR, typename T, typename E>
__device__ __forceinline__ void inject_grad(const E &tb, T *Tx, T *Ty, T *Tz, int l, const T (&td)[2][3])
{
  T *Ts[3] = {Tx, Ty, Tz};
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    T v[N];
    lds_read_line<N, N, DIR>(Ts[e], l, v);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += tb.fv0[i] * td[0][e] + tb.fv1[i] * td[1][e];
    lds_write_line<N, N, DIR>(Ts[e], l, v);
  }
}

template <int N, typename T, bool CHEBY, bool DIAG>
__global__ void __launch_bounds__(kGenMaxThreads) __attribute__((amdgpu_waves_per_eu(1, gen_waves<N, T, DIAG>())))
sipg_general_kernel(const GenArgs<N, T> a)
{
  constexpr int N2 = N * N, N3 = N * N * N, p = N - 1;
  const int CB = a.cb;
  // dynamic LDS: Q [CB][N3], G [3][CB][N3], F [CB][12][N2] (matvec only)
  T *sQ = dyn_lds<T>();
  T *sG[3] = {sQ + CB * N3, sQ + 2 * CB * N3, sQ + 3 * CB * N3};
  T *sF = sQ + 4 * CB * N3;
  const int t = threadIdx.x, c = t / N2, l = t % N2;
  const long long item0 = (long long)logical_block(blockIdx.x, gridDim.x, a.order) * CB, item = item0 + c;
  const bool act = c < CB && item < a.nitems;
  const long long cell = DIAG ? item / N3 : traverse_cell(item, a.nx, a.ny, a.nz, a.order);
  const int unit = DIAG ? (int)(item % N3) : 0; // diagonal: dof of the unit vector
  T *Q = sQ + (act ? c : 0) * N3;
  T *Gx = sG[0] + (act ? c : 0) * N3;
  T *Gy = sG[1] + (act ? c : 0) * N3;
  T *Gz = sG[2] + (act ? c : 0) * N3;
  T *F = sF + (!DIAG && act ? c * 12 * N2 : 0);
  // periodic neighbour cells below / above in x, y, z
  int nbr[6] = {0, 0, 0, 0, 0, 0};
  if (act) periodic_neighbors((int)cell, a.nx, a.ny, a.nz, nbr);

  // ------------------------------------------------------------------ L
  if (act) {
    const auto &tb = ctab(a.dt);
    T u[N], v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if constexpr (DIAG) u[k] = (k * N2 + l == unit) ? T(1) : T(0);
      else u[k] = a.src[cell * N3 + k * N2 + l];
    }
    eo_apply<N, 1>(tb.Sf, u, v);
    lds_write_line<N, N, 2>(Q, l, v);
  }
  if constexpr (!DIAG) {
    // neighbour layers: thread l of a cell loads face point l of all six
    // faces (twelve independent loads, issued together)
    if (act) {
      T uf[6], ui[6];
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const int d = f / 2, s = f % 2;
        const T *nb = a.src + (long long)nbr[f] * N3;
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
  // ----------------------------------------------------------------- V1
  if (act) {
    const auto &tb = ctab(a.dt);
    T u[N], v[N];
    lds_read_line<N, N, 0>(Q, l, u);
    eo_apply<N, 1>(tb.Sf, u, v);
    lds_write_line<N, N, 0>(Q, l, v);
    if constexpr (!DIAG) {
      for (int it = l; it < 12 * N; it += N2) {
        T *fl = F + (it / N) * N2 + (it % N) * N;
#pragma unroll
        for (int i = 0; i < N; ++i) u[i] = fl[i];
        eo_apply<N, 1>(tb.Sf, u, v);
#pragma unroll
        for (int i = 0; i < N; ++i) fl[i] = v[i];
      }
    }
  }
  __syncthreads();
  // ----------------------------------------------------------------- V2
  if (act) {
    const auto &tb = ctab(a.dt);
    T u[N], v[N];
    lds_read_line<N, N, 1>(Q, l, u);
    eo_apply<N, 1>(tb.Sf, u, v);
    lds_write_line<N, N, 1>(Q, l, v);
    if constexpr (!DIAG) {
      for (int it = l; it < 12 * N; it += N2) {
        T *fl = F + (it / N) * N2 + (it % N);
#pragma unroll
        for (int i = 0; i < N; ++i) u[i] = fl[i * N];
        eo_apply<N, 1>(tb.Sf, u, v);
#pragma unroll
        for (int i = 0; i < N; ++i) fl[i * N] = v[i];
      }
    }
  }
  __syncthreads();
  // ------------------------------------------------------------------ G
  if (act) {
    const auto &tb = ctab(a.dt);
    T u[N], v[N];
    lds_read_line<N, N, 0>(Q, l, u);
    eo_apply<N, -1>(tb.Df, u, v);
    lds_write_line<N, N, 0>(Gx, l, v);
    lds_read_line<N, N, 1>(Q, l, u);
    eo_apply<N, -1>(tb.Df, u, v);
    lds_write_line<N, N, 1>(Gy, l, v);
    lds_read_line<N, N, 2>(Q, l, u);
    eo_apply<N, -1>(tb.Df, u, v);
    lds_write_line<N, N, 2>(Gz, l, v);
  }
  __syncthreads();
  // ------------------------------------------------------------------ E
  T tvx[2], tdx[2][3], tvy[2], tdy[2][3], tvz[2], tdz[2][3];
  if (act) {
    const auto &tb = ctab(a.dt);
    // collocation-derivative entries D[m][l % N], D[m][l / N] of this
    // thread's in-face indices (tangential derivatives of the neighbour's
    // face data), the same for all three face directions
    T dlo[N], dhi[N];
    if constexpr (!DIAG) {
      const auto &dtb = cdevtab(a.dt);
#pragma unroll
      for (int m = 0; m < N; ++m) {
        dlo[m] = dtb.Dd[m][l % N];
        dhi[m] = dtb.Dd[m][l / N];
      }
    }
    // (loading the face metric of the next direction ahead, or of all
    // three, measured no faster: the registers cost occupancy)
    T fmv[2][7];
    load_face_metric<N, 0>(a, nbr[0], cell, l, fmv);
    face_terms<N, 0, T, DIAG>(a, Q, Gx, Gy, Gz, F, l, dlo, dhi, fmv, tvx, tdx);
    load_face_metric<N, 1>(a, nbr[2], cell, l, fmv);
    face_terms<N, 1, T, DIAG>(a, Q, Gx, Gy, Gz, F, l, dlo, dhi, fmv, tvy, tdy);
    load_face_metric<N, 2>(a, nbr[4], cell, l, fmv);
    face_terms<N, 2, T, DIAG>(a, Q, Gx, Gy, Gz, F, l, dlo, dhi, fmv, tvz, tdz);
  }
  __syncthreads();
  // ------------------------------------------------------------------ M
  if (act) {
    // the metric of this z line in two batches of 3 N loads, each issued
    // together (two memory round trips per wave instead of N; all 6 N at
    // once would cost a wave per SIMD in registers)
    const T *m = a.cm + cell * 6 * N3 + l;
    T gm[3][N];
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
      for (int k = 0; k < N; ++k) gm[e][k] = load_stream(m + e * N3 + k * N2);
    const auto &tb = ctab(a.dt);
    T gx[N], gy[N], gz[N];
    lds_read_line<N, N, 2>(Gx, l, gx);
    lds_read_line<N, N, 2>(Gy, l, gy);
    lds_read_line<N, N, 2>(Gz, l, gz);
    T hx[N], hy[N], hz[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      hx[k] = gm[0][k] * gx[k];
      hy[k] = gm[1][k] * gy[k];
      hz[k] = gm[2][k] * gz[k];
    }
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
      for (int k = 0; k < N; ++k) gm[e][k] = load_stream(m + (3 + e) * N3 + k * N2);
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const T g01 = gm[0][k], g02 = gm[1][k], g12 = gm[2][k];
      const T x = gx[k], y = gy[k], z = gz[k];
      gx[k] = hx[k] + g01 * y + g02 * z + tb.fv0[k] * tdz[0][0] + tb.fv1[k] * tdz[1][0];
      gy[k] = g01 * x + hy[k] + g12 * z + tb.fv0[k] * tdz[0][1] + tb.fv1[k] * tdz[1][1];
      gz[k] = g02 * x + g12 * y + hz[k] + tb.fv0[k] * tdz[0][2] + tb.fv1[k] * tdz[1][2];
    }
    lds_write_line<N, N, 2>(Gx, l, gx);
    lds_write_line<N, N, 2>(Gy, l, gy);
    lds_write_line<N, N, 2>(Gz, l, gz);
  }
  __syncthreads();
  // ----------------------------------------------------------------- Iy
  if (act) inject_grad<N, 1, T>(ctab(a.dt), Gx, Gy, Gz, l, tdy);
  __syncthreads();
  // ----------------------------------------------------------------- Ix
  if (act) {
    const auto &tb = ctab(a.dt);
    T v[N], r[N];
    lds_read_line<N, N, 0>(Gx, l, v);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += tb.fv0[i] * tdx[0][0] + tb.fv1[i] * tdx[1][0];
    eo_apply<N, -1>(tb.Db, v, r);
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] += tb.fv0[i] * tvx[0] + tb.fv1[i] * tvx[1];
    lds_write_line<N, N, 0>(Q, l, r);
    // tangential x-face tests into T_y, T_z
#pragma unroll
    for (int e = 1; e < 3; ++e) {
      T *Te = e == 1 ? Gy : Gz;
      lds_read_line<N, N, 0>(Te, l, v);
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] += tb.fv0[i] * tdx[0][e] + tb.fv1[i] * tdx[1][e];
      lds_write_line<N, N, 0>(Te, l, v);
    }
  }
  __syncthreads();
  // ----------------------------------------------------------------- By
  if (act) {
    const auto &tb = ctab(a.dt);
    T v[N], r[N], b[N];
    lds_read_line<N, N, 1>(Gy, l, v);
    eo_apply<N, -1>(tb.Db, v, r);
    lds_read_line<N, N, 1>(Q, l, b);
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] += b[i] + tb.fv0[i] * tvy[0] + tb.fv1[i] * tvy[1];
    lds_write_line<N, N, 1>(Q, l, r);
  }
  __syncthreads();
  // ----------------------------------------------------------------- Bz
  if (act) {
    const auto &tb = ctab(a.dt);
    T v[N], r[N], b[N];
    lds_read_line<N, N, 2>(Gz, l, v);
    eo_apply<N, -1>(tb.Db, v, r);
    lds_read_line<N, N, 2>(Q, l, b);
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] += b[i] + tb.fv0[i] * tvz[0] + tb.fv1[i] * tvz[1];
    eo_apply<N, 1>(tb.Sb, r, v);
    lds_write_line<N, N, 2>(Q, l, v);
  }
  __syncthreads();
  // ----------------------------------------------------------------- Sy
  // Chebyshev: rhs, diag, tmp and the input at the dofs this thread stores
  // (its z line, dofs l + m N^2), loaded two phases ahead of the update
  T crhs[CHEBY ? N : 1], cdiag[CHEBY ? N : 1], ctm[CHEBY ? N : 1], csrc[CHEBY ? N : 1];
  if (act) {
    if constexpr (CHEBY) {
      const long long g0 = cell * N3 + l;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const long long g = g0 + i * N2;
        crhs[i] = a.rhs[g];
        cdiag[i] = a.diag[g];
        ctm[i] = a.c_tm != T(0) ? a.tm[g] : T(0); // c_tm == 0: tm not read
        csrc[i] = a.src[g];
      }
    }
    const auto &tb = ctab(a.dt);
    T u[N], v[N];
    lds_read_line<N, N, 1>(Q, l, u);
    eo_apply<N, 1>(tb.Sb, u, v);
    lds_write_line<N, N, 1>(Q, l, v);
  }
  __syncthreads();
  // ----------------------------------------------------------------- Sx
  if (act) {
    const auto &tb = ctab(a.dt);
    T u[N], o[N];
    lds_read_line<N, N, 0>(Q, l, u);
    eo_apply<N, 1>(tb.Sb, u, o);
    if constexpr (DIAG) {
      if (unit / N == l) a.dst[cell * N3 + unit] = o[unit % N];
    } else if constexpr (CHEBY) {
      lds_write_line<N, N, 0>(Q, l, o); // this thread's own line: no race
    } else {
      // written once, read by nobody in this launch: non-temporal stores
      // keep L2 for the neighbour layers and face metrics other workgroups
      // re-read (measured 3-4 % faster on MI355X).  Staged through LDS as
      // z lines like the Chebyshev step below, the extra barrier cost more
      // (150 -> 158 us) than the coalescing saved.
      const long long g0 = cell * N3 + l * N;
#pragma unroll
      for (int i = 0; i < N; ++i) __builtin_nontemporal_store(o[i], a.dst + g0 + i);
    }
  }
  if constexpr (CHEBY && !DIAG) {
    // the Chebyshev step's six streams (rhs, diag, tmp, src in; tmp, dst
    // out) move as z lines (dofs l + m N^2): every access is consecutive
    // across lanes (the Chebyshev(5) sweep at p5 32^3: 1.135 -> 1.04 ms)
    __syncthreads();
    if (act) {
      T o[N];
      lds_read_line<N, N, 2>(Q, l, o);
      const long long g0 = cell * N3 + l;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const long long g = g0 + i * N2;
        const T res = o[i] - crhs[i];
        const T base = a.c_np * cdiag[i] * res;
        const T tmv = a.c_tm != T(0) ? a.c_tm * ctm[i] + base : base;
        a.tm[g] = tmv;
        a.dst[g] = csrc[i] - tmv;
      }
    }
  }
}

template <int N, typename T>
static hipError_t launch_general_n(const SipgLaunch &L, hipStream_t s)
{
  GenArgs<N, T> a{};
  a.dt = (const DevTab<N, T> *)L.dtab;
  a.src = (const T *)L.src;
  a.dst = (T *)L.dst;
  a.rhs = (const T *)L.rhs;
  a.diag = (const T *)L.diag;
  a.tm = (T *)L.tm;
  a.c_np = T(L.c_np);
  a.c_tm = T(L.c_tm);
  a.cm = (const T *)L.cell_metric;
  a.fm = (const T *)L.face_metric;
  a.ncells = L.ncells;
  a.nitems = L.diagonal ? L.ncells * N * N * N : L.ncells;
  a.nx = L.nc[0];
  a.ny = L.nc[1];
  a.nz = L.nc[2];
  a.cb = L.cb > 0 ? L.cb : gen_cells_per_block<N>();
  a.order = L.order;
  if (a.cb * N * N > kGenMaxThreads) return hipErrorInvalidConfiguration;
  const size_t lds = size_t(a.cb) * (4 * N * N * N + (L.diagonal ? 0 : 12 * N * N)) * sizeof(T);
  const dim3 grid((unsigned)((a.nitems + a.cb - 1) / a.cb)), block(((a.cb * N * N + 63) / 64) * 64);
  auto kern = L.diagonal ? sipg_general_kernel<N, T, false, true>
                         : (L.cheby ? sipg_general_kernel<N, T, true, false> : sipg_general_kernel<N, T, false, false>);
  const hipError_t e = allow_dynamic_lds((const void *)kern, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, grid, block, lds, s, a);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatch_general(const SipgLaunch &L, hipStream_t s)
{
  switch (L.tables->n) {
#define SFK_CASE(NN) \
  case NN: return launch_general_n<NN, T>(L, s);
    SFK_CASE(2) SFK_CASE(3) SFK_CASE(4) SFK_CASE(5) SFK_CASE(6) SFK_CASE(7) SFK_CASE(8) SFK_CASE(9) \
    SFK_CASE(10) SFK_CASE(11) SFK_CASE(12) SFK_CASE(13) SFK_CASE(14) SFK_CASE(15) SFK_CASE(16)
#undef SFK_CASE
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sipg_general(const SipgLaunch &L, hipStream_t s)
{
  return L.fp32 ? dispatch_general<float>(L, s) : dispatch_general<double>(L, s);
}

// ------------------------------------------------------------------------
// sine-deformed mesh metric: x = xhat + amp sin(pi xhat) sin(pi yhat)
// sin(pi zhat) (1,1,1) on the uniform mesh of [0,1]^3 (setup, FP64 math)
// ------------------------------------------------------------------------
struct MetricArgs {
  double xg[16], wg[16];
  int n, nx, ny, nz;
  double amp;
  long long ncells;
  double *vol;
  double *area; // [d][cell]
  void *cm, *fm;
  int fp32;
};

__device__ inline void sine_jacobian(const MetricArgs &a, const double xh[3], double J[3][3])
{
  constexpr double pi = 3.14159265358979323846;
  const double sx = sin(pi * xh[0]), sy = sin(pi * xh[1]), sz = sin(pi * xh[2]);
  const double cx = cos(pi * xh[0]), cy = cos(pi * xh[1]), cz = cos(pi * xh[2]);
  const double g[3] = {a.amp * pi * cx * sy * sz, a.amp * pi * sx * cy * sz, a.amp * pi * sx * sy * cz};
  const double h[3] = {1.0 / a.nx, 1.0 / a.ny, 1.0 / a.nz};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[i][j] = h[j] * ((i == j ? 1.0 : 0.0) + g[j]);
}

__device__ inline double inverse3(const double J[3][3], double Ji[3][3])
{
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double id = 1.0 / det;
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  return det;
}

__device__ inline void store(void *base, long long idx, double v, int fp32)
{
  if (fp32) ((float *)base)[idx] = (float)v;
  else ((double *)base)[idx] = v;
}

__device__ inline void cell_coords(const MetricArgs &a, long long c, int ci[3])
{
  ci[0] = (int)(c % a.nx);
  const long long r = c / a.nx;
  ci[1] = (int)(r % a.ny);
  ci[2] = (int)(r / a.ny);
}

__global__ void metric_cell_kernel(const MetricArgs a)
{
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= a.ncells) return;
  const int n = a.n, n3 = n * n * n;
  int ci[3];
  cell_coords(a, c, ci);
  const double h[3] = {1.0 / a.nx, 1.0 / a.ny, 1.0 / a.nz};
  double vol = 0.;
  for (int q = 0; q < n3; ++q) {
    const int qi[3] = {q % n, (q / n) % n, q / (n * n)};
    double xh[3], J[3][3], Ji[3][3];
    for (int d = 0; d < 3; ++d) xh[d] = (ci[d] + a.xg[qi[d]]) * h[d];
    sine_jacobian(a, xh, J);
    const double det = inverse3(J, Ji);
    const double jxw = det * a.wg[qi[0]] * a.wg[qi[1]] * a.wg[qi[2]];
    vol += jxw;
    double G[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) G[i][j] = jxw * (Ji[i][0] * Ji[j][0] + Ji[i][1] * Ji[j][1] + Ji[i][2] * Ji[j][2]);
    const long long b = c * 6 * n3 + q;
    store(a.cm, b, G[0][0], a.fp32);
    store(a.cm, b + n3, G[1][1], a.fp32);
    store(a.cm, b + 2 * n3, G[2][2], a.fp32);
    store(a.cm, b + 3 * n3, G[0][1], a.fp32);
    store(a.cm, b + 4 * n3, G[0][2], a.fp32);
    store(a.cm, b + 5 * n3, G[1][2], a.fp32);
  }
  if (a.vol) a.vol[c] = vol;
}

// pass 1 (write = 0): face areas; pass 2 (write = 1): face data with tau
__global__ void metric_face_kernel(const MetricArgs a, int write)
{
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (id >= 3 * a.ncells) return;
  const int d = (int)(id / a.ncells);
  const long long c = id % a.ncells;
  const int n = a.n, n2 = n * n, p = n - 1;
  int ci[3], cu[3];
  cell_coords(a, c, ci);
  for (int e = 0; e < 3; ++e) cu[e] = ci[e];
  const int nd = d == 0 ? a.nx : (d == 1 ? a.ny : a.nz);
  cu[d] = (cu[d] + 1) % nd;
  const long long u = ((long long)cu[2] * a.ny + cu[1]) * a.nx + cu[0];
  const int lo = d == 0 ? 1 : 0, hi = d == 2 ? 1 : 2;
  const double h[3] = {1.0 / a.nx, 1.0 / a.ny, 1.0 / a.nz};
  double tau = 0.;
  if (write) {
    const double A = a.area[(long long)d * a.ncells + c];
    tau = double(p + 1) * double(p + 1) * 0.5 * (A / a.vol[c] + A / a.vol[u]);
  }
  double area = 0.;
  for (int fa = 0; fa < n2; ++fa) {
    double xiL[3], xiU[3];
    xiL[lo] = xiU[lo] = a.xg[fa % n];
    xiL[hi] = xiU[hi] = a.xg[fa / n];
    xiL[d] = 1.0;
    xiU[d] = 0.0;
    double xL[3], xU[3], JL[3][3], JU[3][3], JLi[3][3], JUi[3][3];
    for (int e = 0; e < 3; ++e) {
      xL[e] = (ci[e] + xiL[e]) * h[e];
      xU[e] = (cu[e] + xiU[e]) * h[e];
    }
    sine_jacobian(a, xL, JL);
    sine_jacobian(a, xU, JU);
    const double detL = inverse3(JL, JLi);
    inverse3(JU, JUi);
    double nv[3] = {JLi[d][0], JLi[d][1], JLi[d][2]}; // J^{-T} e_d
    const double nrm = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    const double wf = a.wg[fa % n] * a.wg[fa / n];
    const double jxwf = detL * nrm * wf;
    area += jxwf;
    if (!write) continue;
    for (int e = 0; e < 3; ++e) nv[e] /= nrm;
    const long long b = ((long long)d * a.ncells + c) * 7 * n2 + fa;
    for (int e = 0; e < 3; ++e) {
      store(a.fm, b + e * n2, jxwf * (JLi[e][0] * nv[0] + JLi[e][1] * nv[1] + JLi[e][2] * nv[2]), a.fp32);
      store(a.fm, b + (3 + e) * n2, jxwf * (JUi[e][0] * nv[0] + JUi[e][1] * nv[1] + JUi[e][2] * nv[2]), a.fp32);
    }
    store(a.fm, b + 6 * n2, tau * jxwf, a.fp32);
  }
  if (!write) a.area[(long long)d * a.ncells + c] = area;
}

hipError_t launch_sine_metric(const Tables1D &t, bool fp32, const int nc[3], double amp, void *cell_metric,
                              void *face_metric, hipStream_t s)
{
  MetricArgs a{};
  a.n = t.n;
  for (int i = 0; i < t.n; ++i) {
    a.xg[i] = t.gauss[i];
    a.wg[i] = t.weights[i];
  }
  a.nx = nc[0];
  a.ny = nc[1];
  a.nz = nc[2];
  a.amp = amp;
  a.ncells = (long long)nc[0] * nc[1] * nc[2];
  a.cm = cell_metric;
  a.fm = face_metric;
  a.fp32 = fp32 ? 1 : 0;
  hipError_t e = hipMalloc(&a.vol, sizeof(double) * a.ncells * 4);
  if (e != hipSuccess) return e;
  a.area = a.vol + a.ncells;
  const int bs = 128;
  hipLaunchKernelGGL(metric_cell_kernel, dim3((unsigned)((a.ncells + bs - 1) / bs)), dim3(bs), 0, s, a);
  if (face_metric) {
    const unsigned g = (unsigned)((3 * a.ncells + bs - 1) / bs);
    hipLaunchKernelGGL(metric_face_kernel, dim3(g), dim3(bs), 0, s, a, 0);
    hipLaunchKernelGGL(metric_face_kernel, dim3(g), dim3(bs), 0, s, a, 1);
  }
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(a.vol);
  return e;
}

} // namespace sfk
