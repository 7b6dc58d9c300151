"""Writes the golden vectors under tests/golden/ from the reference's own C++
code (oracle/_ref/libppref.so, compiled from the reference sources with strict
IEEE flags and no OpenMP, so every result is deterministic).

Run in the container that holds /root/reference:
    make -C oracle ref && python tests/golden/make_golden.py

Inputs come from the library's seeded generators (SplitMix64; see oracle/ppo.c),
so a machine without the reference can regenerate them.  Matrices come from
the reference's generateLaplacianCSR where it has one; the upwind operator has
no reference generator and is produced by the oracle, then handed to the
reference's ILUT/GMRES as a CSR.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import ppo  # noqa: E402


def save(name, **arrays):
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print("wrote", name)


def main():
    R = ppo.Ref("strict")

    # --- generators + SpMV / SpMV^T --------------------------------------
    for dims, conv in [((8, 8, 1), (0.0, 0.0, 0.0, 0.0)), ((6, 5, 4), (0.2, 0.1, 0.05, 0.3)),
                       ((7, 1, 1), (0.0, 0.0, 0.0, 0.0))]:
        A = R.laplacian(*dims, *conv)
        n = len(A[0]) - 1
        x = ppo.uniform(n, 7)
        tag = "x".join(map(str, dims))
        save(f"spmv_{tag}", dims=np.array(dims), conv=np.array(conv), rp=A[0], ci=A[1], v=A[2],
             x=x, y=R.spmv(A, x), yt=R.spmv_t(A, x))

    # --- threshold IC and LDL^T -----------------------------------------
    for dims in [(8, 8, 1), (16, 16, 1), (6, 5, 4)]:
        A = R.laplacian(*dims)
        n = len(A[0]) - 1
        hA = R.matrix(A)
        x = ppo.uniform(n, 11)
        tag = "x".join(map(str, dims))
        for tau in (0.0, 1e-3, 1e-2):
            hp = R.prec(hA, "ict", tau)
            lt = R.factor(hp, 0, n)
            save(f"ict_{tag}_tau{tau:g}", dims=np.array(dims), tau=tau, lt_rp=lt[0], lt_ci=lt[1],
                 lt_v=lt[2], mnnz=R.L.refh_prec_mnnz(hp), x=x, apply=R.apply(hp, x),
                 lower=R.apply(hp, x, 0), upper=R.apply(hp, x, 2))
            R.L.refh_prec_free(hp)
            hp = R.prec(hA, "ildlt", tau)
            save(f"ildlt_{tag}_tau{tau:g}", dims=np.array(dims), tau=tau, x=x,
                 mnnz=R.L.refh_prec_mnnz(hp), apply=R.apply(hp, x), lower=R.apply(hp, x, 0),
                 diag=R.apply(hp, x, 1), upper=R.apply(hp, x, 2))
            R.L.refh_prec_free(hp)
        R.free_matrix(hA)

    # --- ILUT on nonsymmetric operators ---------------------------------
    cases = [("lapconv_6x5x4", R.laplacian(6, 5, 4, 0.2, 0.1, 0.05, 0.0)),
             ("upwind_8x8x8", ppo.upwind(8, 8, 8, 20.0, 10.0, 5.0)),
             ("lapconv_16x16x1", R.laplacian(16, 16, 1, 0.3, 0.2, 0.0, 0.0))]
    for tag, A in cases:
        n = len(A[0]) - 1
        hA = R.matrix(A)
        x = ppo.uniform(n, 13)
        for lfil, tau in ((20, 1e-2), (3, 0.0), (5, 1e-3)):
            hp = R.prec(hA, "ilut", tau, lfil)
            L, U, D = R.factor(hp, 1, n), R.factor(hp, 2, n), R.ilu_diag(hp, n)
            save(f"ilut_{tag}_l{lfil}_tau{tau:g}", rp=A[0], ci=A[1], v=A[2], lfil=lfil, tau=tau,
                 l_rp=L[0], l_ci=L[1], l_v=L[2], u_rp=U[0], u_ci=U[1], u_v=U[2], d=D,
                 mnnz=R.L.refh_prec_mnnz(hp), x=x, apply=R.apply(hp, x))
            R.L.refh_prec_free(hp)
        R.free_matrix(hA)

    # --- FlexibleCG + ICT (config 1 and smaller) ------------------------
    for dims, tau in [((128, 128, 1), 1e-3), ((128, 128, 1), 1e-2), ((16, 16, 1), 1e-3),
                      ((12, 12, 12), 1e-3), ((20, 20, 20), 1e-2)]:
        A = R.laplacian(*dims)
        n = len(A[0]) - 1
        xt = ppo.uniform(n, 0)
        b = R.spmv(A, xt)
        hA = R.matrix(A)
        hp = R.prec(hA, "ict", tau)
        out = R.cg(hA, hp, b, tol=1e-8, maxit=1000)
        out_n = R.cg(hA, hp, b, tol=1e-8, maxit=1000, flexible=False)
        tag = "x".join(map(str, dims))
        save(f"cg_{tag}_tau{tau:g}", dims=np.array(dims), tau=tau, seed=0, tol=1e-8, maxit=1000,
             iters=out["iters"], converged=out["converged"], error=out["error"], x=out["x"],
             normal_iters=out_n["iters"], normal_error=out_n["error"])
        R.L.refh_prec_free(hp)
        R.free_matrix(hA)

    # --- FlexibleGMRES(30) + ILUT ---------------------------------------
    for tag, A, lfil, tau in [("lapconv_12x12x12", R.laplacian(12, 12, 12, 0.2, 0.1, 0.05, 0.0), 20, 1e-2),
                              ("upwind_16x16x16", ppo.upwind(16, 16, 16, 20.0, 10.0, 5.0), 20, 1e-2),
                              ("upwind_10x10x10_l3", ppo.upwind(10, 10, 10, 20.0, 10.0, 5.0), 3, 0.0)]:
        n = len(A[0]) - 1
        xt = ppo.uniform(n, 0)
        b = R.spmv(A, xt)
        hA = R.matrix(A)
        hp = R.prec(hA, "ilut", tau, lfil)
        out = R.gmres(hA, hp, b, tol=1e-8, restart=30)
        save(f"gmres_{tag}", rp=A[0], ci=A[1], v=A[2], lfil=lfil, tau=tau, seed=0, tol=1e-8,
             restart=30, iters=out["iters"], converged=out["converged"], error=out["error"], x=out["x"])
        R.L.refh_prec_free(hp)
        R.free_matrix(hA)


if __name__ == "__main__":
    main()
