"""Write the golden fixtures of tests/golden/*.npz from the REFERENCE
implementation, compiled from the reference headers by oracle/Makefile into
oracle/_ref/libsfkref.so (this container only; the reference does not travel).

    make -C oracle ref && python tests/golden/make_golden.py [--high]

(--high: only the fixtures of the degrees above 8)

Inputs are regenerated from the stored seeds (numpy default_rng); only the
reference outputs are stored, in the canonical layout (cell-major, x
fastest), converted from the reference's lane-interleaved layout.  Cells
whose +x neighbour the reference reads from a wrong address (global x index
nx-1, legacy/include/evaluation_dg_laplacian.h:270) are recorded in `mask`
(1 = valid).  Degree 2 of the Hermite-basis operators is left out: the
reference's backward symmetric 1D sweep drops the middle coefficient
(matrix_vector_kernel.h:130-131, type 0), which is exact only when the basis
is nodal at x = 1/2 (GLL, Hermite p >= 3 odd / p >= 4 even) and not for the
p = 2 Hermite basis.  PrecomputedDGLaplacian stores every degree: its
outputs at the rows that the reference computes from wrong addresses are
masked by the tests that read them (tests/test_oracle_precomputed.py).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bindings as ob  # noqa: E402

CELL_IJ = (4.0, 2.0, 4.0 / 3.0)  # evaluation_cell_laplacian.h:803-813 / evaluation_dg_laplacian.h:605-613


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"{name}: {os.path.getsize(path) / 1024:.1f} KiB")


def main():
    if not ob.ref_available():
        raise SystemExit("oracle/_ref/libsfkref.so missing: run `make -C oracle ref`")
    R = ob.ref()
    L = ob.ref_lanes()

    # quadrature and bases
    q = {}
    for n in range(1, 13):
        for fn, key in ((R.ref_gauss_points, "gauss_points"), (R.ref_gauss_weights, "gauss_weights")):
            a = np.zeros(n)
            fn(n, ob._ptr(a))
            q[f"{key}_{n}"] = a
        if n >= 2:
            a = np.zeros(n)
            R.ref_gll_points(n, ob._ptr(a))
            q[f"gll_points_{n}"] = a
    xs = np.linspace(0.0, 1.0, 11)
    for p in range(1, 10):
        q[f"hermite_value_{p}"] = np.array([[R.ref_hermite_value(p, i, x) for x in xs] for i in range(p + 1)])
        q[f"hermite_derivative_{p}"] = np.array([[R.ref_hermite_derivative(p, i, x) for x in xs]
                                                 for i in range(p + 1)])
    q["hermite_x"] = xs
    save("quadrature_bases.npz", **q)

    t = {}
    for kind in (0, 1, 2):
        for p in range(1, 9):
            tb = ob.ref_tables(kind, p)
            for k, v in tb.items():
                t[f"{kind}_{p}_{k}"] = np.atleast_1d(v)
    save("tables.npz", **t)

    # 1D kernel: (nn, type, contract_over_rows, add_into, do_dg)
    k1 = {}
    variants = [(0, 1, 0, 0), (0, 0, 0, 0), (1, 1, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (0, 0, 1, 0),
                (1, 1, 0, 2), (1, 0, 0, 2), (1, 0, 1, 2)]
    for nn in range(2, 10):
        tb = ob.ref_tables(1, nn - 1)
        for vi, (typ, cor, add, dg) in enumerate(variants):
            for stride in (1, nn):
                seed = 1000 + 100 * nn + 10 * vi + stride
                rng = np.random.default_rng(seed)
                ln = stride * (nn - 1) + 1
                inp, out0, addv, face = (rng.standard_normal(ln), rng.standard_normal(ln), rng.standard_normal(ln),
                                         rng.standard_normal(4))
                coeffs = tb["values_eo"] if typ == 0 else tb["gradients_eo"]
                out, fc = out0.copy(), face.copy()
                rc = R.ref_apply_1d(nn, typ, cor, add, dg, stride, ob._ptr(coeffs), ob._ptr(inp), ob._ptr(out),
                                    ob._ptr(addv), ob._ptr(tb["face_eo"]), ob._ptr(fc))
                assert rc == 0
                k1[f"{nn}_{vi}_{stride}_out"] = out
                k1[f"{nn}_{vi}_{stride}_face"] = fc
    k1["variants"] = np.array(variants)
    save("kernel1d.npz", **k1)

    # element-local cell Laplacian (Hermite basis), READ_VECTOR build
    for p in [1, 3, 4, 5, 6, 7, 8]:
        n3 = (p + 1) ** 3
        nb = 3
        nc = (nb * L, 1, 1)
        seed = 10 + p
        src = np.random.default_rng(seed).standard_normal(nb * L * n3)  # canonical
        out_i = ob.ref_cell_laplacian(p, ob.to_interleaved(p, nc, L, src), CELL_IJ)
        save(f"cell_laplacian_p{p}.npz", seed=seed, n_cells=np.array(nc), inv_jac=np.array(CELL_IJ),
             out=ob.to_canonical(p, nc, L, out_i))

    # DG-SIPG matvec and Chebyshev step
    for p in [1, 3, 4, 5, 6, 7, 8]:
        nc = (2 * L, 2, 3) if p <= 5 else (L, 2, 2)
        n3 = (p + 1) ** 3
        N = n3 * int(np.prod(nc))
        seed = 20 + p
        rng = np.random.default_rng(seed)
        src, rhs, diag, tm = rng.standard_normal(N), rng.standard_normal(N), rng.uniform(0.5, 1.5, N), \
            rng.standard_normal(N)
        I = lambda x: ob.to_interleaved(p, nc, L, x)  # noqa: E731
        C = lambda x: ob.to_canonical(p, nc, L, x)  # noqa: E731
        mv, _ = ob.ref_dg_laplacian(p, nc, CELL_IJ, I(src))
        new, tmo = ob.ref_dg_laplacian(p, nc, CELL_IJ, I(src), I(rhs), I(diag), I(tm), 0.7, 0.3, mode=1)
        mask = np.ones((nc[2], nc[1], nc[0]), dtype=np.int8)
        mask[:, :, nc[0] - 1] = 0
        save(f"dg_laplacian_p{p}.npz", seed=seed, n_cells=np.array(nc), inv_jac=np.array(CELL_IJ), c_np=0.7,
             c_tm=0.3, matvec=C(mv), cheby_new=C(new), cheby_tm=C(tmo), mask=mask.ravel())

    # DG advection matvec and RK stage (reference geometry: [0,2]^3)
    for p in range(1, 9):
        nc = (2 * L, 2, 3) if p <= 4 else (L, 2, 2)
        n3 = (p + 1) ** 3
        N = n3 * int(np.prod(nc))
        seed = 30 + p
        rng = np.random.default_rng(seed)
        src, upd, tm = rng.standard_normal(N), rng.standard_normal(N), rng.standard_normal(N)
        I = lambda x: ob.to_interleaved(p, nc, L, x)  # noqa: E731
        C = lambda x: ob.to_canonical(p, nc, L, x)  # noqa: E731
        mv, _ = ob.ref_dg_advect(p, nc, None, I(src))
        vnp, vtm = ob.ref_dg_advect(p, nc, None, I(src), I(upd), I(tm), 0.3, 0.2, mode=1)
        save(f"dg_advect_p{p}.npz", seed=seed, n_cells=np.array(nc),
             inv_jac=np.array([nc[0] / 2.0, nc[1] / 2.0, nc[2] / 2.0]), velocity=np.array([1.0, 0.15, -0.05]),
             c_np=0.3, c_tm=0.2, matvec=C(mv), stage_np=C(vnp), stage_tm=C(vtm))

    # PrecomputedDGLaplacian (precomputed_dg_laplacian.h), unit cells
    for p in range(1, 9):
        nc = (2 * L, 3, 2) if p <= 5 else (L, 2, 2)
        n3 = (p + 1) ** 3
        N = n3 * int(np.prod(nc))
        seed = 50 + p
        src = np.random.default_rng(seed).standard_normal(N)
        out, m_eo, l_eo, outer = ob.ref_dg_precomputed(p, nc, ob.to_interleaved(p, nc, L, src))
        save(f"dg_precomputed_p{p}.npz", seed=seed, n_cells=np.array(nc), matvec=ob.to_canonical(p, nc, L, out),
             mass_eo=m_eo, laplace_eo=l_eo, value_outer=np.array(outer))

    # advection known-answer tests (evaluation_dg_advect.h:221-283)
    kat = {}
    for p in range(1, 8):
        nc = (2 * L, 8, 8)
        ncu = (ob.ctypes.c_uint * 3)(*nc)
        ed, es, te, ij = np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(3)
        n_steps = 20
        rc = R.ref_advect_kat(p, ncu, n_steps, ob._ptr(ed), ob._ptr(es), ob._ptr(te), ob._ptr(ij))
        assert rc == 0
        kat[f"p{p}"] = np.array([ed[0], es[0], te[0], n_steps])
        kat[f"inv_jac_p{p}"] = ij
    kat["n_cells"] = np.array((2 * L, 8, 8))
    save("advect_kat.npz", **kat)


def high_degrees():
    """fixtures of the degrees above 8 (the reference drivers run the DG
    Laplacian to degree 12, advection to 14 and the cell Laplacian beyond):
    one x row of L cells each, to keep the files small"""
    if not ob.ref_available():
        raise SystemExit("oracle/_ref/libsfkref.so missing: run `make -C oracle ref`")
    R = ob.ref()
    L = ob.ref_lanes()
    nc = (L, 1, 1)
    for p in (10, 12, 15):
        n3 = (p + 1) ** 3
        seed = 10 + p
        src = np.random.default_rng(seed).standard_normal(L * n3)
        out_i = ob.ref_cell_laplacian(p, ob.to_interleaved(p, nc, L, src), CELL_IJ)
        save(f"cell_laplacian_p{p}.npz", seed=seed, n_cells=np.array(nc), inv_jac=np.array(CELL_IJ),
             out=ob.to_canonical(p, nc, L, out_i))
    for p in (10, 12):
        N = (p + 1) ** 3 * L
        seed = 20 + p
        rng = np.random.default_rng(seed)
        src, rhs, diag, tm = rng.standard_normal(N), rng.standard_normal(N), rng.uniform(0.5, 1.5, N), \
            rng.standard_normal(N)
        I = lambda x: ob.to_interleaved(p, nc, L, x)  # noqa: E731
        C = lambda x: ob.to_canonical(p, nc, L, x)  # noqa: E731
        mv, _ = ob.ref_dg_laplacian(p, nc, CELL_IJ, I(src))
        new, tmo = ob.ref_dg_laplacian(p, nc, CELL_IJ, I(src), I(rhs), I(diag), I(tm), 0.7, 0.3, mode=1)
        mask = np.ones((nc[2], nc[1], nc[0]), dtype=np.int8)
        mask[:, :, nc[0] - 1] = 0
        save(f"dg_laplacian_p{p}.npz", seed=seed, n_cells=np.array(nc), inv_jac=np.array(CELL_IJ), c_np=0.7,
             c_tm=0.3, matvec=C(mv), cheby_new=C(new), cheby_tm=C(tmo), mask=mask.ravel())
    for p in (11, 14):
        N = (p + 1) ** 3 * L
        seed = 30 + p
        rng = np.random.default_rng(seed)
        src, upd, tm = rng.standard_normal(N), rng.standard_normal(N), rng.standard_normal(N)
        I = lambda x: ob.to_interleaved(p, nc, L, x)  # noqa: E731
        C = lambda x: ob.to_canonical(p, nc, L, x)  # noqa: E731
        mv, _ = ob.ref_dg_advect(p, nc, None, I(src))
        vnp, vtm = ob.ref_dg_advect(p, nc, None, I(src), I(upd), I(tm), 0.3, 0.2, mode=1)
        save(f"dg_advect_p{p}.npz", seed=seed, n_cells=np.array(nc),
             inv_jac=np.array([nc[0] / 2.0, nc[1] / 2.0, nc[2] / 2.0]), velocity=np.array([1.0, 0.15, -0.05]),
             c_np=0.3, c_tm=0.2, matvec=C(mv), stage_np=C(vnp), stage_tm=C(vtm))
    kat = {}
    nck = (L, 4, 4)
    for p in (10, 14):
        ncu = (ob.ctypes.c_uint * 3)(*nck)
        ed, es, te, ij = np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(3)
        n_steps = 10
        rc = R.ref_advect_kat(p, ncu, n_steps, ob._ptr(ed), ob._ptr(es), ob._ptr(te), ob._ptr(ij))
        assert rc == 0
        kat[f"p{p}"] = np.array([ed[0], es[0], te[0], n_steps])
        kat[f"inv_jac_p{p}"] = ij
    kat["n_cells"] = np.array(nck)
    save("advect_kat_high.npz", **kat)


if __name__ == "__main__":
    if sys.argv[1:] == ["--high"]:
        high_degrees()
    else:
        main()
        high_degrees()
