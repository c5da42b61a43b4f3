"""Helpers shared by the GPU tests: run the same node batch through the HIP
kernel (C-ABI, torch device memory) and through the CPU oracle."""
import ctypes

import numpy as np

import gpuboscia_amd as gb


def torch_mod():
    import torch
    return torch


def dev(a, dtype=None):
    torch = torch_mod()
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a.copy())
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


class Problem:
    """A dense-quadratic or QAP objective plus an LMO, held on both sides."""

    def __init__(self, kind="dense", Q=None, c=None, gs=2.0, ls=0.5, kappa=0.0, A=None, B=None,
                 lmo="box", rhs=1.0, equality=False, is_int=None, order=0, cap=0.2, card=10.0):
        from oracle import pyoracle as orc
        self.orc = orc
        self.kind, self.lmo_kind = kind, lmo
        self.keep = []
        if kind == "dense":
            Q = np.ascontiguousarray(Q, dtype=np.float64)
            self.n = Q.shape[0]
            Qt = np.ascontiguousarray(Q.T)
            self.dQt = dev(Qt)
            self.dc = None if c is None else dev(np.asarray(c, dtype=np.float64))
            self.gobj = gb.ObjectiveDesc(gb.OBJ_DENSE, self.n, self.dQt.data_ptr(),
                                         0 if self.dc is None else self.dc.data_ptr(), gs, ls, kappa,
                                         0, 0, 0)
            self.oobj = orc.dense_objective(Q, c, gs, ls, kappa)
        else:
            A = np.ascontiguousarray(A, dtype=np.float64)
            B = np.ascontiguousarray(B, dtype=np.float64)
            m = A.shape[0]
            self.n = m * m
            self.dA, self.dB = dev(A), dev(B)
            self.gobj = gb.ObjectiveDesc(gb.OBJ_QAP, self.n, 0, 0, 1.0, 1.0, 0.0, m,
                                         self.dA.data_ptr(), self.dB.data_ptr())
            self.oobj = orc.qap_objective(A, B)
        n = self.n
        if lmo == "box":
            self.glmo = gb.LmoDesc(gb.LMO_BOX, 1.0, 0, 0, 0)
            self.olmo = orc.box_lmo()
        elif lmo == "simplex":
            ii = np.ascontiguousarray(is_int if is_int is not None else np.zeros(n), dtype=np.uint8)
            self.dii = dev(ii)
            self.glmo = gb.LmoDesc(gb.LMO_SIMPLEX, float(rhs), 1 if equality else 0,
                                   self.dii.data_ptr(), 0)
            self.olmo = orc.simplex_lmo(rhs, ii, equality)
        elif lmo == "portfolio":
            m = n // 2
            self.glmo = gb.LmoDesc(gb.LMO_PORTFOLIO, float(rhs), 0, 0, 0, float(cap), float(card))
            self.olmo = orc.portfolio_lmo(m, cap, card, rhs)
        else:
            self.glmo = gb.LmoDesc(gb.LMO_BIRKHOFF, 1.0, 0, 0, int(order))
            self.olmo = orc.birkhoff_lmo(order)

    def run_gpu(self, alg, X0, LB, UB, EPS, ITERS, STATUS, flags, max_iters, T=0, ctol=1e-6):
        torch = torch_mod()
        N, n = X0.shape
        x, lb, ub = dev(X0), dev(LB), dev(UB)
        eps = dev(EPS)
        it = dev(ITERS.astype(np.int32))
        st = dev(STATUS.astype(np.uint8))
        gap = torch.zeros(N, dtype=torch.float64, device="cuda")
        val = torch.zeros(N, dtype=torch.float64, device="cuda")
        bv = torch.zeros((N, n), dtype=torch.float64, device="cuda")
        bvv = torch.full((N,), 1e300, dtype=torch.float64, device="cuda")
        fl = None if flags is None else dev(np.asarray(flags, dtype=np.uint8))
        gb.relax_batch(self.gobj, alg_gpu(alg), self.glmo, N, max_iters, x, lb, ub, eps, it, st,
                       gap, val, bv, bvv, fl, ctol, T)
        torch.cuda.synchronize()
        return {"x": x.cpu().numpy(), "gap": gap.cpu().numpy(), "value": val.cpu().numpy(),
                "iterations": it.cpu().numpy(), "status": st.cpu().numpy(),
                "best_vertex": bv.cpu().numpy(), "best_vertex_val": bvv.cpu().numpy()}

    def run_orc(self, alg, X0, LB, UB, EPS, ITERS, STATUS, flags, max_iters, T=64, ctol=1e-6,
                threads=8):
        r = self.orc.relax_batch(self.oobj, alg_orc(alg), self.olmo, np.array(X0, dtype=np.float64),
                                 np.ascontiguousarray(LB, dtype=np.float64),
                                 np.ascontiguousarray(UB, dtype=np.float64),
                                 np.ascontiguousarray(EPS, dtype=np.float64),
                                 np.array(ITERS, dtype=np.int32), np.array(STATUS, dtype=np.uint8),
                                 flags=flags, max_iters=max_iters, candidate_tol=ctol, lanes=T,
                                 threads=threads)
        return r


def alg_gpu(a):
    return gb.algorithm(dicg=a["dicg"], line_search=a["ls"], heuristic=a.get("heur", 0),
                        heuristic_offset=a.get("off", 0.0))


def alg_orc(a):
    from oracle import pyoracle as orc
    return orc.alg(dicg=a["dicg"], ls=a["ls"], heuristic=a.get("heur", 0),
                   heuristic_offset=a.get("off", 0.0))


def assert_bit_equal(g, o, label=""):
    for k in ("x", "gap", "value", "iterations", "status", "best_vertex", "best_vertex_val"):
        a, b = np.asarray(g[k]), np.asarray(o[k])
        if a.dtype.kind == "f":
            same = (a == b) | (np.isnan(a) & np.isnan(b))
            if not np.all(same):
                idx = np.argwhere(~same)[0]
                raise AssertionError(f"{label} {k} differs at {tuple(idx)}: gpu={a[tuple(idx)]!r} "
                                     f"oracle={b[tuple(idx)]!r}")
        else:
            assert np.array_equal(a, b), f"{label} {k}: gpu={a} oracle={b}"


def auto_T(n):
    w = ((n + 63) // 64) * 64
    return max(64, min(256, w))
