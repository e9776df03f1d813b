// DG Laplacian with precomputed 1D matrices on Cartesian cells
// (legacy/include/precomputed_dg_laplacian.h:96-377, do_inner_loop):
//
//   A u = f_x (M (x) M (x) L~) u + f_y (M (x) L~ (x) M) u + f_z (L~ (x) M (x) M) u
//
// (x fastest on the right), L = 1D Laplace matrix with the cell's own
// interior-penalty face terms, M = 1D mass matrix, both symmetric and
// centro-symmetric in the Hermite-like basis; L~ adds the coupling of rows
// 0, 1 (n-1, n-2) to the lower (upper) neighbour's two face layers through
// the constants v1, v2 (:139-148, :175-188, :215-224).  f_d =
// inv_jac_d^2 / (inv_jac_x inv_jac_y inv_jac_z) (1 on the reference's unit
// cells).  Periodic in every direction.
//
// One workgroup = cb consecutive cells, n^2 threads per cell; thread l owns
// line l of every direction.  Phases (barrier between):
//   0  cell block -> LDS U (coalesced), neighbour layers -> registers
//   1  p = L~_x u (x line) -> P, q = L~_y u (y line) -> Q, r = L~_z u (z line, registers)
//   2  P <- M_z P, Q <- M_z Q (z lines), U <- r
//   3  P <- M_y P, Q <- Q + M_y U (y lines)
//   4  U <- M_x Q + P (x lines)
//   5  U -> dst (coalesced), or the fused Chebyshev update of
//      evaluation_dg_laplacian.h:485-502 on the way out.
// The same order of sums as the reference: f_x M_z M_y L~_x + M_x (M_z
// f_y L~_y + M_y f_z L~_z).
#include "sfk_device.h"
#include "sfk_kernels.h"

namespace sfk {

namespace {

constexpr int kPcMaxThreads = 256;

template <int N, typename T>
struct PcTab {
  Eo<N, T> L, M;  // symmetric: the same table serves both contraction directions
  T v1, v2;       // neighbour coupling
  T f[3];         // direction factors
  T ld[N], md[N]; // diagonals of L and M (operator diagonal)
};

template <int N, typename T>
__device__ __forceinline__ const SFK_CONST_AS PcTab<N, T> &pctab(const PcTab<N, T> *p)
{
  return *(const SFK_CONST_AS PcTab<N, T> *)opaque_ptr(p);
}

template <int N>
constexpr int pc_cells_per_block()
{
  return kPcMaxThreads / (N * N) > 0 ? kPcMaxThreads / (N * N) : 1;
}

template <int N, typename T>
constexpr int pc_waves()
{
  constexpr int cb = pc_cells_per_block<N>();
  return lds_waves_per_simd(3 * cb * CellLds<N, T>::CS * (int)sizeof(T), cb * N * N);
}

template <int N, typename T>
struct PcArgs {
  const PcTab<N, T> *pt;
  const T *src;
  T *dst;
  const T *rhs, *diag;
  T *tm;
  T c_np, c_tm;
  long long ncells, nvec;
  int nx, ny, nz;
  int cb, order;
};

// index inside a cell of point k of line l in direction DIR (x fastest)
template <int N, int DIR>
__device__ __forceinline__ int pc_idx(int l, int k)
{
  return line_base<N, N, DIR>(l) + k * line_stride<N, N, DIR>();
}

// rows 0, 1 couple to the lower neighbour's layers n-1, n-2; rows n-1, n-2
// to the upper neighbour's layers 0, 1 (same adds as the reference)
template <int N, typename T>
__device__ __forceinline__ void pc_couple(T (&o)[N], const T (&nb)[4], T v1, T v2)
{
  o[0] += v1 * nb[0] + v2 * nb[1];
  o[1] += v2 * nb[0];
  o[N - 1] += v1 * nb[2] + v2 * nb[3];
  o[N - 2] += v2 * nb[2];
}

template <int N, int DIR, typename T>
__device__ __forceinline__ void pc_load_layers(const T *src, const int (&nbr)[6], int l, T (&nb)[4])
{
  constexpr long long N3 = N * N * N;
  const T *lo = src + nbr[2 * DIR] * N3;
  const T *hi = src + nbr[2 * DIR + 1] * N3;
  nb[0] = lo[pc_idx<N, DIR>(l, N - 1)];
  nb[1] = lo[pc_idx<N, DIR>(l, N - 2)];
  nb[2] = hi[pc_idx<N, DIR>(l, 0)];
  nb[3] = hi[pc_idx<N, DIR>(l, 1)];
}

template <int N, typename T, bool COUPLE, bool CHEBY>
__global__ void __launch_bounds__(kPcMaxThreads) __attribute__((amdgpu_waves_per_eu(1, pc_waves<N, T>())))
precomp_kernel(PcArgs<N, T> a)
{
  constexpr int N2 = N * N, N3 = N2 * N;
  const int cb = a.cb;
  constexpr int CS = CellLds<N, T>::CS;
  T *U = dyn_lds<T>();
  T *P = U + cb * CS, *Q = P + cb * CS;
  const long long c0 = (long long)logical_block(blockIdx.x, gridDim.x, a.order) * cb;
  const long long voff = (long long)blockIdx.y * a.nvec;
  const T *src = a.src + voff;
  const int t = threadIdx.x, c = t / N2, l = t - c * N2;
  const long long cell = c0 + c;
  const bool act = c < cb && cell < a.ncells;
  const long long rem = a.ncells - c0;
  const int nval = (rem < cb ? (int)rem : cb) * N3;
  const T *sb = src + c0 * N3;

  T nb[3][4];
  if (COUPLE && act) {
    int nbr[6];
    periodic_neighbors((int)cell, a.nx, a.ny, a.nz, nbr);
    pc_load_layers<N, 0>(src, nbr, l, nb[0]);
    pc_load_layers<N, 1>(src, nbr, l, nb[1]);
    pc_load_layers<N, 2>(src, nbr, l, nb[2]);
  }
  // blockDim.x >= cb n^2, so n rounds cover the cb n^3 values of the block;
  // the loads are branch-free (index clamped into the block) so that all n
  // are in flight together, and only the LDS writes are guarded
  T keep[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int i = t + k * (int)blockDim.x;
    keep[k] = sb[i < nval ? i : nval - 1];
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int i = t + k * (int)blockDim.x;
    if (i < nval) U[(i / N3) * CS + cl_off<N, T>(i % N3)] = keep[k];
  }
  __syncthreads();

  T *Uc = U + c * CS, *Pc = P + c * CS, *Qc = Q + c * CS;
  T r[N];
  if (act) {
    const auto &pt = pctab(a.pt);
    T v[N], o[N];
    cl_read<N, 0>(Uc, l, v);
    eo_apply<N, 1>(pt.L, v, o);
    if constexpr (COUPLE) pc_couple(o, nb[0], pt.v1, pt.v2);
#pragma unroll
    for (int i = 0; i < N; ++i) o[i] *= pt.f[0];
    cl_write<N, 0>(Pc, l, o);
    cl_read<N, 1>(Uc, l, v);
    eo_apply<N, 1>(pt.L, v, o);
    if constexpr (COUPLE) pc_couple(o, nb[1], pt.v1, pt.v2);
#pragma unroll
    for (int i = 0; i < N; ++i) o[i] *= pt.f[1];
    cl_write<N, 1>(Qc, l, o);
    cl_read<N, 2>(Uc, l, v);
    eo_apply<N, 1>(pt.L, v, r);
    if constexpr (COUPLE) pc_couple(r, nb[2], pt.v1, pt.v2);
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] *= pt.f[2];
  }
  __syncthreads();
  if (act) {
    const auto &pt = pctab(a.pt);
    T v[N], o[N];
    cl_read<N, 2>(Pc, l, v);
    eo_apply<N, 1>(pt.M, v, o);
    cl_write<N, 2>(Pc, l, o);
    cl_read<N, 2>(Qc, l, v);
    eo_apply<N, 1>(pt.M, v, o);
    cl_write<N, 2>(Qc, l, o);
    cl_write<N, 2>(Uc, l, r);
  }
  __syncthreads();
  if (act) {
    const auto &pt = pctab(a.pt);
    T v[N], o[N];
    cl_read<N, 1>(Pc, l, v);
    eo_apply<N, 1>(pt.M, v, o);
    cl_write<N, 1>(Pc, l, o);
    cl_read<N, 1>(Uc, l, v);
    eo_apply<N, 1>(pt.M, v, o);
    cl_read<N, 1>(Qc, l, v);
#pragma unroll
    for (int i = 0; i < N; ++i) o[i] = v[i] + o[i];
    cl_write<N, 1>(Qc, l, o);
  }
  __syncthreads();
  if (act) {
    const auto &pt = pctab(a.pt);
    T v[N], o[N];
    cl_read<N, 0>(Qc, l, v);
    eo_apply<N, 1>(pt.M, v, o);
    cl_read<N, 0>(Pc, l, v);
#pragma unroll
    for (int i = 0; i < N; ++i) o[i] = o[i] + v[i];
    cl_write<N, 0>(Uc, l, o);
  }
  __syncthreads();
  const long long g0 = voff + c0 * N3;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int i = t + k * (int)blockDim.x;
    if (i < nval) {
      const T au = U[(i / N3) * CS + cl_off<N, T>(i % N3)];
      if constexpr (CHEBY) {
        // r = A u - rhs; tm = c_tm tm + c_np diag r; dst = u - tm (tm not
        // read when c_tm == 0)
        const long long gi = g0 + i;
        const T res = au - a.rhs[gi];
        T tmv = a.c_np * a.diag[gi] * res;
        if (a.c_tm != T(0)) tmv = a.c_tm * a.tm[gi] + tmv;
        a.tm[gi] = tmv;
        a.dst[gi] = keep[k] - tmv;
      } else {
        __builtin_nontemporal_store(au, a.dst + g0 + i); // written once
      }
    }
  }
}

// diag(A) = f_x L_ii M_jj M_kk + f_y M_ii L_jj M_kk + f_z M_ii M_jj L_kk
// (the coupling is off-diagonal), the same pattern in every cell
template <int N, typename T>
__global__ void __launch_bounds__(256) precomp_diag_kernel(const PcTab<N, T> *ptp, long long n, T *diag)
{
  constexpr int N3 = N * N * N;
  const auto &pt = pctab(ptp);
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long long)gridDim.x * blockDim.x) {
    const int d = (int)(g % N3);
    const int i = d % N, j = (d / N) % N, k = d / (N * N);
    diag[g] = pt.f[0] * pt.ld[i] * pt.md[j] * pt.md[k] + pt.f[1] * pt.md[i] * pt.ld[j] * pt.md[k] +
              pt.f[2] * pt.md[i] * pt.md[j] * pt.ld[k];
  }
}

template <int N, typename T>
hipError_t launch_precomp_n(const PrecompLaunch &L, hipStream_t s)
{
  PcArgs<N, T> a{};
  a.pt = (const PcTab<N, T> *)L.ptab;
  a.src = (const T *)L.src;
  a.dst = (T *)L.dst;
  a.rhs = (const T *)L.rhs;
  a.diag = (const T *)L.diag;
  a.tm = (T *)L.tm;
  a.c_np = T(L.c_np);
  a.c_tm = T(L.c_tm);
  a.ncells = L.ncells;
  a.nvec = L.ncells * N * N * N;
  a.nx = L.nc[0];
  a.ny = L.nc[1];
  a.nz = L.nc[2];
  const int CB = L.cb > 0 ? L.cb : pc_cells_per_block<N>();
  if (CB * N * N > kPcMaxThreads) return hipErrorInvalidConfiguration;
  if (L.cheby && L.nrhs != 1) return hipErrorInvalidValue;
  a.cb = CB;
  a.order = L.order;
  const size_t lds = size_t(3) * CB * CellLds<N, T>::CS * sizeof(T);
  const dim3 grid((unsigned)((L.ncells + CB - 1) / CB), (unsigned)L.nrhs), block(((CB * N * N + 63) / 64) * 64);
  decltype(&precomp_kernel<N, T, true, false>) kern;
  if (L.couple) kern = L.cheby ? precomp_kernel<N, T, true, true> : precomp_kernel<N, T, true, false>;
  else kern = L.cheby ? precomp_kernel<N, T, false, true> : precomp_kernel<N, T, false, false>;
  const hipError_t e = allow_dynamic_lds((const void *)kern, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, grid, block, lds, s, a);
  return hipGetLastError();
}

template <int N, typename T>
size_t precomp_lds_per_cell_n()
{
  return size_t(3) * CellLds<N, T>::CS * sizeof(T);
}

template <int N, typename T>
hipError_t launch_precomp_diag_n(const PrecompLaunch &L, hipStream_t s)
{
  const long long n = L.ncells * N * N * N;
  const long long blocks = (n + 255) / 256;
  const unsigned g = (unsigned)(blocks < 65536 ? blocks : 65536);
  hipLaunchKernelGGL((precomp_diag_kernel<N, T>), dim3(g), dim3(256), 0, s, (const PcTab<N, T> *)L.ptab, n,
                     (T *)L.dst);
  return hipGetLastError();
}

template <int N, typename T>
hipError_t upload_pctab(const Precomp1D &pc, const double f[3], void **dev)
{
  PcTab<N, T> h{};
  h.L = make_eo<N, T>(pc.lap, true);
  h.M = make_eo<N, T>(pc.mass, true);
  h.v1 = T(pc.v1);
  h.v2 = T(pc.v2);
  for (int d = 0; d < 3; ++d) h.f[d] = T(f[d]);
  for (int i = 0; i < N; ++i) {
    h.ld[i] = T(pc.lap[i * N + i]);
    h.md[i] = T(pc.mass[i * N + i]);
  }
  hipError_t e = hipMalloc(dev, sizeof(h));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dev, &h, sizeof(h), hipMemcpyHostToDevice);
}

} // namespace

#define SFK_PC_DISPATCH(fn, n, ...)                                                                \
  switch (n) {                                                                                     \
  case 2: return fn<2, T>(__VA_ARGS__);                                                            \
  case 3: return fn<3, T>(__VA_ARGS__);                                                            \
  case 4: return fn<4, T>(__VA_ARGS__);                                                            \
  case 5: return fn<5, T>(__VA_ARGS__);                                                            \
  case 6: return fn<6, T>(__VA_ARGS__);                                                            \
  case 7: return fn<7, T>(__VA_ARGS__);                                                            \
  case 8: return fn<8, T>(__VA_ARGS__);                                                            \
  case 9: return fn<9, T>(__VA_ARGS__);                                                            \
  default: return hipErrorInvalidValue;                                                            \
  }

template <typename T>
static hipError_t dispatch_precomp(const PrecompLaunch &L, hipStream_t s)
{
  SFK_PC_DISPATCH(launch_precomp_n, L.n, L, s)
}
template <typename T>
static hipError_t dispatch_precomp_diag(const PrecompLaunch &L, hipStream_t s)
{
  SFK_PC_DISPATCH(launch_precomp_diag_n, L.n, L, s)
}
template <typename T>
static hipError_t dispatch_pctab(int n, const Precomp1D &pc, const double f[3], void **dev)
{
  SFK_PC_DISPATCH(upload_pctab, n, pc, f, dev)
}
#undef SFK_PC_DISPATCH

hipError_t launch_precomp(const PrecompLaunch &L, hipStream_t s)
{
  return L.fp32 ? dispatch_precomp<float>(L, s) : dispatch_precomp<double>(L, s);
}

hipError_t launch_precomp_diagonal(const PrecompLaunch &L, hipStream_t s)
{
  return L.fp32 ? dispatch_precomp_diag<float>(L, s) : dispatch_precomp_diag<double>(L, s);
}

hipError_t upload_precomp_tables(int n, const Precomp1D &pc, const double f[3], bool fp32, void **dev)
{
  return fp32 ? dispatch_pctab<float>(n, pc, f, dev) : dispatch_pctab<double>(n, pc, f, dev);
}

size_t precomp_lds_per_cell(int n, bool fp32)
{
  switch (n) {
#define SFK_CASE(NN) \
  case NN: return fp32 ? precomp_lds_per_cell_n<NN, float>() : precomp_lds_per_cell_n<NN, double>();
    SFK_CASE(2) SFK_CASE(3) SFK_CASE(4) SFK_CASE(5) SFK_CASE(6) SFK_CASE(7) SFK_CASE(8) SFK_CASE(9)
#undef SFK_CASE
  default: return size_t(3) * n * n * n * (fp32 ? 4 : 8);
  }
}

} // namespace sfk
