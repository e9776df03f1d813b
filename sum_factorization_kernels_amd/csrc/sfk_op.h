// The operator handle behind the opaque sfk_op of include/sfk.h.
#pragma once

#include "sfk.h"
#include "sfk_internal.h"
#include "sfk_kernels.h"

struct sfk_op {
  sfk_op_desc d{};
  sfk::Tables1D tab;
  long long ncells = 0;
  long long nvec = 0; // elements of one vector
  long long ncq = 0;  // cell quadrature points (all cells)
  long long nfq = 0;  // face quadrature points per direction (all cells)
  int grid[3] = {0, 0, 0};
  // device data owned by the handle (handle precision)
  void *cell_metric = nullptr, *face_metric = nullptr, *coef = nullptr;
  void *cell_vel = nullptr, *face_vel = nullptr;
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  void *dtab = nullptr; // 1D tables in device memory (uploaded on first launch)
  // workgroup schedule (sfk_schedule): cells per block (0 = default), block
  // order, continuous scatter (-1 automatic, 0 atomic, 1 colored, 2
  // tensor-product kernel) and the tensor-product kernel's tile shape and
  // cells in z per workgroup (0 = automatic)
  int sched_cb = 0, sched_order = 1, cont_scatter = -1;
  int tp_shape = 0, tp_zc = 0;
  // the tensor-product kernel's Chebyshev launches (sfk_cheby_step, the
  // solver's smoother): tile shape, cells in z and block order chosen by
  // sfk_tune for those launches (cheby_shape -1: the matvec's)
  int cheby_shape = -1, cheby_zc = 0, cheby_order = 1;
  // the per-q-metric SIPG kernel's fused Chebyshev step: cells per block
  // chosen by sfk_tune for it, with cheby_order (0: the matvec's schedule)
  int cheby_cb = 0;
  void *tptab = nullptr; // TpTab of the tensor-product kernel (constant coefficient)
  // precomputed DG Laplacian: neighbour coupling constants not both zero
  // (set with the device tables)
  bool pc_couple = true;
};

namespace sfk {

inline void fill_sipg_launch(const sfk_op *op, SipgLaunch &L, bool cheby = false)
{
  L.tables = &op->tab;
  L.fp32 = op->d.precision == SFK_FP32;
  L.ncells = op->ncells;
  for (int d = 0; d < 3; ++d) L.nc[d] = op->d.n_cells[d];
  L.cell_metric = op->cell_metric;
  L.face_metric = op->face_metric;
  L.dtab = op->dtab;
  L.cb = op->sched_cb;
  L.order = op->sched_order;
  if (cheby && op->cheby_cb > 0) {
    L.cb = op->cheby_cb;
    L.order = op->cheby_order;
  }
}

inline void fill_advect_launch(const sfk_op *op, AdvectLaunch &L)
{
  L.tables = &op->tab;
  L.fp32 = op->d.precision == SFK_FP32;
  L.ncells = op->ncells;
  for (int d = 0; d < 3; ++d) {
    L.nc[d] = op->d.n_cells[d];
    L.inv_jac[d] = op->d.inv_jac[d];
    L.vel[d] = op->d.velocity[d];
  }
  if (op->d.flags & SFK_FLAG_VARIABLE_VELOCITY) {
    L.cell_vel = op->cell_vel;
    L.face_vel = op->face_vel;
  }
  L.dtab = op->dtab;
  L.cb = op->sched_cb;
  L.order = op->sched_order;
}

inline void fill_precomp_launch(const sfk_op *op, PrecompLaunch &L)
{
  L.n = op->d.degree + 1;
  L.fp32 = op->d.precision == SFK_FP32;
  L.couple = op->pc_couple;
  L.ncells = op->ncells;
  for (int d = 0; d < 3; ++d) L.nc[d] = op->d.n_cells[d];
  L.nrhs = op->d.n_rhs;
  L.ptab = op->dtab;
  L.cb = op->sched_cb;
  L.order = op->sched_order;
}

size_t elem_size(const sfk_op *op);
// device tables of the operator (uploaded on first use)
sfk_status ensure_device_tables(sfk_op *op);
// the continuous Laplacian's launches use the tensor-product kernel
bool uses_tp(const sfk_op *op);
// (cheby: the schedule of the fused Chebyshev launches)
void fill_tp_launch(const sfk_op *op, TpLaunch &L, bool cheby = false);

} // namespace sfk
