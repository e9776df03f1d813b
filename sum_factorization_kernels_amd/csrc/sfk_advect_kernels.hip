// DG upwind advection on a Cartesian periodic mesh (GLL nodal basis,
// Gauss quadrature), EvaluationDGAdvection::do_inner_loop / do_rk_stage of
// legacy/include/evaluation_dg_advect.h:285-715, plus what the reference
// does not have: per-point velocity, the adjoint (transposed) operator and
// the gradient with respect to the velocity.
//
// kind 0  r = M^{-1} L(u),  L(u)_K = (grad v, a u)_K - <v, (a.n) u_upwind>;
//         optional fused LSRK stage np = upd + c_np r, tm = np + c_tm r.
// kind 1  L^T y:  (v, a.grad y)_K - <v, max(a.n,0) (y^- - y^+)>  (downwind
//         coupling), so that A^T w = L^T (M^{-1} w).
// kind 2  M^{-1} x = Sinv^T W^{-1} Sinv x (cell-local collocation inverse).
// kind 3  velocity gradient of mu^T L(a) u with mu = M^{-1} lambda:
//         cell: (d mu/dx_d) u JxW at every q-point; face (upper face of each
//         cell, normal velocity b = a.n):  -(mu^- - mu^+) u_upwind JxW_f.
// Thread/phase organisation as in sfk_cell_kernels.hip.
#include "sfk_device.h"
#include "sfk_kernels.h"

namespace sfk {

template <int N>
constexpr int adv_cells_per_block()
{
  return 256 / (N * N) > 0 ? 256 / (N * N) : 1;
}
constexpr int kAdvMaxThreads = 256;
template <int N, typename T, int KIND>
constexpr int adv_waves()
{
  constexpr int cb = adv_cells_per_block<N>();
  return lds_waves_per_simd(cb * (2 * CellLds<N, T>::CS + (KIND == 2 ? 0 : 6 * N * N)) * (int)sizeof(T), cb * N * N);
}

template <int N, typename T>
struct AdvArgs {
  const DevTab<N, T> *dt; // tables in device memory (read through ctab)
  const T *src, *upd, *aux;
  T *dst, *tm, *dst2;
  T c_np, c_tm;
  const T *cell_vel, *face_vel;
  T vel[3];
  T ij[3];
  T jxw;
  T jxwf[3];
  long long ncells;
  int nx, ny, nz;
  int cb;    // cells per workgroup
  int order; // block order (logical_block)
};

// normal velocity a.n (outward from the cell) at face point fa of face (d, s)
// (the lower face's value is stored with the lower neighbour `lower`)
template <int N, typename T, bool VAR>
__device__ __forceinline__ T normal_velocity(const AdvArgs<N, T> &a, long long cell, int lower, int d, int s, int fa)
{
  constexpr int N2 = N * N;
  if constexpr (!VAR) {
    return s ? a.vel[d] : -a.vel[d];
  } else {
    if (s) return a.face_vel[((long long)d * a.ncells + cell) * N2 + fa];
    return -a.face_vel[((long long)d * a.ncells + lower) * N2 + fa];
  }
}

template <int N, int DIR, typename T, bool VAR>
__device__ __forceinline__ T cell_velocity(const AdvArgs<N, T> &a, long long cell, int l, int i)
{
  if constexpr (!VAR) return a.vel[DIR];
  else {
    constexpr int N3 = N * N * N;
    const int dof = line_base<N, N, DIR>(l) + i * line_stride<N, N, DIR>();
    return a.cell_vel[(cell * 3 + DIR) * N3 + dof];
  }
}

// value face layers of the 6 neighbours (GLL nodal: the face layer is the
// trace) into this cell's F[f][fa]: thread l loads face point l of every
// face (six independent loads, issued together)
template <int N, typename T>
__device__ __forceinline__ void load_face_layers(const T *src, T *F, const int (&nbr)[6], int l)
{
  constexpr int N2 = N * N, N3 = N * N * N, p = N - 1;
  T v[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2, lf = f % 2 ? 0 : p;
    const int of = d == 0 ? face_dof<N, 0>(lf, l) : (d == 1 ? face_dof<N, 1>(lf, l) : face_dof<N, 2>(lf, l));
    v[f] = src[(long long)nbr[f] * N3 + of];
  }
#pragma unroll
  for (int f = 0; f < 6; ++f) F[f * N2 + l] = v[f];
}

template <int N, typename T, typename E>
__device__ __forceinline__ void face_sweep_t(const E &M, T *F, int narr, int l, bool lo)
{
  constexpr int N2 = N * N;
  for (int it = l; it < narr * N; it += N2) {
    T u[N], v[N];
    T *fl = F + (it / N) * N2 + (lo ? (it % N) * N : (it % N));
    const int st = lo ? 1 : N;
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = fl[i * st];
    eo_apply<N, 1>(M, u, v);
#pragma unroll
    for (int i = 0; i < N; ++i) fl[i * st] = v[i];
  }
}

// forward (kind 0) and transposed weak form (kind 1): direction phase
template <int N, int DIR, typename T, bool VAR, bool ADJ>
__device__ __forceinline__ void adv_dir(const AdvArgs<N, T> &a, const T *Acell, const T *F, int l, long long cell,
                                        int lower, T (&r)[N])
{
  const auto &tb = ctab(a.dt);
  T q[N];
  cl_read<N, DIR>(Acell, l, q);
  const T c = a.ij[DIR] * a.jxw;
  if constexpr (!ADJ) {
    T g[N];
#pragma unroll
    for (int i = 0; i < N; ++i)
      g[i] = cell_velocity<N, DIR, T, VAR>(a, cell, l, i) * c * (line_weight<N, DIR>(tb.w, l, i) * q[i]);
    eo_apply<N, -1>(tb.Db, g, r);
  } else {
    T g[N];
    eo_apply<N, -1>(tb.Df, q, g);
#pragma unroll
    for (int i = 0; i < N; ++i)
      r[i] = cell_velocity<N, DIR, T, VAR>(a, cell, l, i) * c * (line_weight<N, DIR>(tb.w, l, i) * g[i]);
  }
  constexpr int N2 = N * N;
  const T wf = tb.w[l % N] * tb.w[l / N] * a.jxwf[DIR];
  T v0 = 0, v1 = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    v0 += tb.fv0[i] * q[i];
    v1 += tb.fv1[i] * q[i];
  }
  T fl[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const T um = s ? v1 : v0;
    const T up = F[(2 * DIR + s) * N2 + l];
    const T an = normal_velocity<N, T, VAR>(a, cell, lower, DIR, s, l);
    if constexpr (!ADJ) {
      // -0.5((u^- + u^+) a.n + |a.n| (u^- - u^+)), evaluation_dg_advect.h:403-404
      fl[s] = T(-0.5) * ((um + up) * an + fabs(an) * (um - up)) * wf;
    } else {
      fl[s] = -(an > T(0) ? an : T(0)) * (um - up) * wf;
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) r[i] += tb.fv0[i] * fl[0] + tb.fv1[i] * fl[1];
}

template <int N, typename T, int KIND, bool VAR, bool STAGE>
__global__ void __launch_bounds__(kAdvMaxThreads) __attribute__((amdgpu_waves_per_eu(1, adv_waves<N, T, KIND>())))
advect_kernel(const AdvArgs<N, T> a)
{
  constexpr int N2 = N * N, N3 = N * N * N;
  constexpr bool ADJ = KIND == 1;
  const int CB = a.cb;
  // dynamic LDS: A [CB][CS], B [CB][CS] (padded cell blocks), F [CB][6][N2]
  // (not for M^{-1})
  constexpr int CS = CellLds<N, T>::CS;
  T *sA = dyn_lds<T>();
  T *sB = sA + CB * CS;
  T *sF = sB + CB * CS;
  const int t = threadIdx.x, c = t / N2, l = t % N2;
  const long long cell0 = (long long)logical_block(blockIdx.x, gridDim.x, a.order) * CB;
  const bool act = c < CB && cell0 + c < a.ncells;
  // the cell at this traversal position and its periodic neighbour cells
  int nbr[6] = {0, 0, 0, 0, 0, 0};
  const long long cell = act ? locate_cell((unsigned)cell0, (unsigned)c, a.nx, a.ny, a.nz, a.order, nbr) : 0;
  T *A = sA + (act ? c : 0) * CS;
  T *B = sB + (act ? c : 0) * CS;
  T *F = sF + (KIND != 2 && act ? c * 6 * N2 : 0);

  if constexpr (KIND == 2) {
    // M^{-1} x = Sinv^T W^{-1} Sinv x
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
#pragma unroll
      for (int k = 0; k < N; ++k) u[k] = a.src[cell * N3 + k * N2 + l];
      eo_apply<N, 1>(tb.Ib, u, v);
      cl_write<N, 2>(A, l, v);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      cl_read<N, 0>(A, l, u);
      eo_apply<N, 1>(tb.Ib, u, v);
      cl_write<N, 0>(A, l, v);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      cl_read<N, 1>(A, l, u);
      eo_apply<N, 1>(tb.Ib, u, v);
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] /= a.jxw * line_weight<N, 1>(tb.w, l, i);
      eo_apply<N, 1>(tb.If, v, u);
      cl_write<N, 1>(A, l, u);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      cl_read<N, 0>(A, l, u);
      eo_apply<N, 1>(tb.If, u, v);
      cl_write<N, 0>(A, l, v);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      cl_read<N, 2>(A, l, u);
      eo_apply<N, 1>(tb.If, u, v);
#pragma unroll
      for (int k = 0; k < N; ++k) a.dst[cell * N3 + k * N2 + l] = v[k];
    }
    return;
  } else if constexpr (KIND == 3) {
    // velocity gradient: A <- u at q-points, B <- mu at q-points,
    // F[0..2] <- u^+ on the upper faces, F[3..5] <- mu^+ on the upper faces
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
#pragma unroll
      for (int k = 0; k < N; ++k) u[k] = a.src[cell * N3 + k * N2 + l];
      eo_apply<N, 1>(tb.Sf, u, v);
      cl_write<N, 2>(A, l, v);
#pragma unroll
      for (int k = 0; k < N; ++k) u[k] = a.aux[cell * N3 + k * N2 + l];
      eo_apply<N, 1>(tb.Sf, u, v);
      cl_write<N, 2>(B, l, v);
    }
    if (act) {
      T v[6];
#pragma unroll
      for (int fi = 0; fi < 6; ++fi) {
        const int d = fi % 3;
        const T *vec = fi < 3 ? a.src : a.aux;
        const int of = d == 0 ? face_dof<N, 0>(0, l) : (d == 1 ? face_dof<N, 1>(0, l) : face_dof<N, 2>(0, l));
        v[fi] = vec[(long long)nbr[2 * d + 1] * N3 + of];
      }
#pragma unroll
      for (int fi = 0; fi < 6; ++fi) F[fi * N2 + l] = v[fi];
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      cl_read<N, 0>(A, l, u);
      eo_apply<N, 1>(tb.Sf, u, v);
      cl_write<N, 0>(A, l, v);
      cl_read<N, 0>(B, l, u);
      eo_apply<N, 1>(tb.Sf, u, v);
      cl_write<N, 0>(B, l, v);
      face_sweep_t<N, T>(tb.Sf, F, 6, l, true);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      cl_read<N, 1>(A, l, u);
      eo_apply<N, 1>(tb.Sf, u, v);
      cl_write<N, 1>(A, l, v);
      cl_read<N, 1>(B, l, u);
      eo_apply<N, 1>(tb.Sf, u, v);
      cl_write<N, 1>(B, l, v);
      face_sweep_t<N, T>(tb.Sf, F, 6, l, false);
    }
    __syncthreads();
    // cell gradients of line l in direction DIR: the x and y lines go
    // through LDS below, so that every store writes consecutive dofs across
    // lanes (FP64 p3 48^3: 168 -> 151 us)
    T gx[N], gy[N];
    if (act) {
      const auto &tb = ctab(a.dt);
      auto do_dir = [&](auto dirc) {
        constexpr int DIR = decltype(dirc)::value;
        T uq[N], mq[N], g[N];
        cl_read<N, DIR>(A, l, uq);
        cl_read<N, DIR>(B, l, mq);
        eo_apply<N, -1>(tb.Df, mq, g);
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const T gi = g[i] * a.ij[DIR] * uq[i] * (a.jxw * line_weight<N, DIR>(tb.w, l, i));
          if constexpr (DIR == 0) gx[i] = gi;
          else if constexpr (DIR == 1) gy[i] = gi;
          else a.dst[(cell * 3 + DIR) * N3 + line_base<N, N, DIR>(l) + i * line_stride<N, N, DIR>()] = gi;
        }
        T um = 0, mm = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          um += tb.fv1[i] * uq[i];
          mm += tb.fv1[i] * mq[i];
        }
        const T up = F[DIR * N2 + l], mp = F[(3 + DIR) * N2 + l];
        const T b = a.face_vel[((long long)DIR * a.ncells + cell) * N2 + l];
        const T sg = b > T(0) ? T(1) : (b < T(0) ? T(-1) : T(0));
        const T uh = T(0.5) * (um + up) + T(0.5) * sg * (um - up);
        a.dst2[((long long)DIR * a.ncells + cell) * N2 + l] =
            -(mm - mp) * uh * (tb.w[l % N] * tb.w[l / N] * a.jxwf[DIR]);
      };
      do_dir(std::integral_constant<int, 0>());
      do_dir(std::integral_constant<int, 1>());
      do_dir(std::integral_constant<int, 2>());
    }
    // x and y gradients: written into the cell blocks (their q-point values
    // are no longer needed) and stored as z lines, lane l at dof l + m N^2
    __syncthreads();
    if (act) {
      cl_write<N, 0>(A, l, gx);
      cl_write<N, 1>(B, l, gy);
    }
    __syncthreads();
    if (act) {
      T v[N];
      cl_read<N, 2>(A, l, v);
#pragma unroll
      for (int m = 0; m < N; ++m) a.dst[(cell * 3 + 0) * N3 + l + m * N2] = v[m];
      cl_read<N, 2>(B, l, v);
#pragma unroll
      for (int m = 0; m < N; ++m) a.dst[(cell * 3 + 1) * N3 + l + m * N2] = v[m];
    }
    return;
  } else {
    // ------------------------------------------------ L: z values + faces
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
#pragma unroll
      for (int k = 0; k < N; ++k) u[k] = a.src[cell * N3 + k * N2 + l];
      eo_apply<N, 1>(tb.Sf, u, v);
      cl_write<N, 2>(A, l, v);
    }
    if (act) load_face_layers<N, T>(a.src, F, nbr, l);
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      cl_read<N, 0>(A, l, u);
      eo_apply<N, 1>(tb.Sf, u, v);
      cl_write<N, 0>(A, l, v);
      face_sweep_t<N, T>(tb.Sf, F, 6, l, true);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      cl_read<N, 1>(A, l, u);
      eo_apply<N, 1>(tb.Sf, u, v);
      cl_write<N, 1>(A, l, v);
      face_sweep_t<N, T>(tb.Sf, F, 6, l, false);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T r[N];
      adv_dir<N, 0, T, VAR, ADJ>(a, A, F, l, cell, nbr[0], r);
      cl_write<N, 0>(B, l, r);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T r[N], b[N];
      adv_dir<N, 1, T, VAR, ADJ>(a, A, F, l, cell, nbr[2], r);
      cl_read<N, 1>(B, l, b);
#pragma unroll
      for (int i = 0; i < N; ++i) r[i] += b[i];
      cl_write<N, 1>(B, l, r);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T r[N], b[N], o[N];
      adv_dir<N, 2, T, VAR, ADJ>(a, A, F, l, cell, nbr[4], r);
      cl_read<N, 2>(B, l, b);
#pragma unroll
      for (int i = 0; i < N; ++i) r[i] += b[i];
      if constexpr (!ADJ) {
        // inverse mass matrix by collocation, evaluation_dg_advect.h:619-624
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] /= a.jxw * line_weight<N, 2>(tb.w, l, i);
        eo_apply<N, 1>(tb.If, r, o);
      } else {
        eo_apply<N, 1>(tb.Sb, r, o);
      }
      cl_write<N, 2>(A, l, o);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], v[N];
      cl_read<N, 1>(A, l, u);
      if constexpr (!ADJ) eo_apply<N, 1>(tb.If, u, v);
      else eo_apply<N, 1>(tb.Sb, u, v);
      cl_write<N, 1>(A, l, v);
    }
    __syncthreads();
    if (act) {
      const auto &tb = ctab(a.dt);
      T u[N], o[N];
      cl_read<N, 0>(A, l, u);
      if constexpr (!ADJ) eo_apply<N, 1>(tb.If, u, o);
      else eo_apply<N, 1>(tb.Sb, u, o);
      cl_write<N, 0>(B, l, o);
    }
    // the x lines go back through LDS and leave as z lines: lane l holds
    // dofs l + m N^2, so every store writes consecutive dofs across lanes
    __syncthreads();
    if (act) {
      T o[N];
      cl_read<N, 2>(B, l, o);
      const long long g0 = cell * N3 + l;
      if constexpr (STAGE) {
        // evaluation_dg_advect.h:651-667
#pragma unroll
        for (int m = 0; m < N; ++m) {
          const long long g = g0 + m * N2;
          const T vnp = a.upd[g] + a.c_np * o[m];
          a.dst[g] = vnp;
          if (a.c_tm != T(0)) a.tm[g] = vnp + a.c_tm * o[m];
        }
      } else {
        // written once, read by nobody in this launch: non-temporal stores
        // keep L2 for the velocity and neighbour layers other workgroups
        // re-read (forward p = 3, 48^3: 73 -> 60 us on MI355X)
#pragma unroll
        for (int m = 0; m < N; ++m) __builtin_nontemporal_store(o[m], a.dst + g0 + m * N2);
      }
    }
  }
}


template <int N, typename T>
static size_t advect_lds_per_cell_n(bool minv)
{
  return size_t(2 * CellLds<N, T>::CS + (minv ? 0 : 6 * N * N)) * sizeof(T);
}

template <int N, typename T>
static hipError_t launch_advect_n(const AdvectLaunch &L, hipStream_t s)
{
  AdvArgs<N, T> a{};
  a.dt = (const DevTab<N, T> *)L.dtab;
  a.src = (const T *)L.src;
  a.upd = (const T *)L.upd;
  a.aux = (const T *)L.aux;
  a.dst = (T *)L.dst;
  a.tm = (T *)L.tm;
  a.dst2 = (T *)L.dst2;
  a.c_np = T(L.c_np);
  a.c_tm = T(L.c_tm);
  a.cell_vel = (const T *)L.cell_vel;
  a.face_vel = (const T *)L.face_vel;
  const double jxw = 1. / (L.inv_jac[0] * L.inv_jac[1] * L.inv_jac[2]);
  for (int d = 0; d < 3; ++d) {
    a.vel[d] = T(L.vel[d]);
    a.ij[d] = T(L.inv_jac[d]);
    a.jxwf[d] = T(jxw * L.inv_jac[d]);
  }
  a.jxw = T(jxw);
  a.ncells = L.ncells;
  a.nx = L.nc[0];
  a.ny = L.nc[1];
  a.nz = L.nc[2];
  const int CB = L.cb > 0 ? L.cb : adv_cells_per_block<N>();
  if (CB * N * N > kAdvMaxThreads) return hipErrorInvalidConfiguration;
  a.cb = CB;
  a.order = L.order;
  const size_t lds = size_t(CB) * advect_lds_per_cell_n<N, T>(L.kind == 2);
  const dim3 grid((unsigned)((L.ncells + CB - 1) / CB)), block(((CB * N * N + 63) / 64) * 64);
  const bool var = L.cell_vel != nullptr;
  decltype(&advect_kernel<N, T, 0, false, false>) kern = nullptr;
  switch (L.kind) {
  case 0:
    if (var) kern = L.stage ? advect_kernel<N, T, 0, true, true> : advect_kernel<N, T, 0, true, false>;
    else kern = L.stage ? advect_kernel<N, T, 0, false, true> : advect_kernel<N, T, 0, false, false>;
    break;
  case 1: kern = var ? advect_kernel<N, T, 1, true, false> : advect_kernel<N, T, 1, false, false>; break;
  case 2: kern = advect_kernel<N, T, 2, false, false>; break;
  case 3: kern = advect_kernel<N, T, 3, true, false>; break;
  default: return hipErrorInvalidValue;
  }
  const hipError_t e = allow_dynamic_lds((const void *)kern, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, grid, block, lds, s, a);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatch_advect(const AdvectLaunch &L, hipStream_t s)
{
  switch (L.tables->n) {
#define SFK_CASE(NN) \
  case NN: return launch_advect_n<NN, T>(L, s);
    SFK_CASE(2) SFK_CASE(3) SFK_CASE(4) SFK_CASE(5) SFK_CASE(6) SFK_CASE(7) SFK_CASE(8) SFK_CASE(9) \
    SFK_CASE(10) SFK_CASE(11) SFK_CASE(12) SFK_CASE(13) SFK_CASE(14) SFK_CASE(15) SFK_CASE(16)
#undef SFK_CASE
  default: return hipErrorInvalidValue;
  }
}

size_t advect_lds_per_cell(int n, bool fp32)
{
  switch (n) {
#define SFK_CASE(NN) \
  case NN: return fp32 ? advect_lds_per_cell_n<NN, float>(false) : advect_lds_per_cell_n<NN, double>(false);
    SFK_CASE(2) SFK_CASE(3) SFK_CASE(4) SFK_CASE(5) SFK_CASE(6) SFK_CASE(7) SFK_CASE(8) SFK_CASE(9) \
    SFK_CASE(10) SFK_CASE(11) SFK_CASE(12) SFK_CASE(13) SFK_CASE(14) SFK_CASE(15) SFK_CASE(16)
#undef SFK_CASE
  default: return size_t(2 * n * n * n + 6 * n * n) * (fp32 ? 4 : 8);
  }
}

hipError_t launch_advect(const AdvectLaunch &L, hipStream_t s)
{
  return L.fp32 ? dispatch_advect<float>(L, s) : dispatch_advect<double>(L, s);
}

} // namespace sfk
