"""Python handle on a batch of periodic cells living on one MI355X.

``CellBatch`` owns one library context (include/maple_hip.h): the multigrid
hierarchy, CG vectors and element energies of every cell, allocated once.
Inputs and outputs are torch tensors on the context's device (float64,
contiguous); every call is enqueued on torch's current stream of that device.
There is no CPU path: a missing library or a missing GPU raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from . import streams as _streams

PHYS = {"elastic": L.MAPLE_ELASTIC, "thermal": L.MAPLE_THERMAL}
NCASE = {"elastic": 6, "thermal": 3}
NDOF = {"elastic": 3, "thermal": 1}


def hex8_matrices(r: int, E0: float = 1.0, nu: float = 0.3, kappa0: float = 1.0):
    """Element matrices of the reference (_utils.py:39-196), from the library's host code.

    Returns numpy float64 (Ke24, Fe24, X0_24, Ke8, Fe8, X0_8).
    """
    lib = L.load()
    out = [np.zeros(s) for s in ((24, 24), (24, 6), (24, 6), (8, 8), (8, 3), (8, 3))]
    L.check(lib.maple_hex8_matrices(int(r), float(E0), float(nu), float(kappa0), *[L.dptr(a) for a in out]))
    return tuple(out)


def max_levels(r: int, min_coarse: int = 4) -> int:
    """Largest level count whose coarsest grid is at least min_coarse per side."""
    n = 1
    while r % (2 ** n) == 0 and r // (2 ** n) >= min_coarse:
        n += 1
    return n


class CellBatch:
    def __init__(self, dims: Sequence, nlevels: int = 4, device=None, precision: str = "fp64",
                 omega: Optional[float] = None, nsmooth: int = 2, warm_start: bool = False,
                 smoother: str = "l1-jacobi", block_cg: bool = False, islands: bool = False):
        self.lib = L.load()
        dl = []
        for d in dims:
            if isinstance(d, (int, np.integer)):
                dl.append((int(d),) * 3)
            else:
                dl.append(tuple(int(v) for v in d))
        self.dims = dl
        self.ncells = len(dl)
        self.nlevels = int(nlevels)
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("CellBatch needs a CUDA (HIP) device; there is no CPU path")
        prec = {"fp64": L.MAPLE_PREC_FP64, "mixed": L.MAPLE_PREC_MIXED}[precision]
        self.precision = precision
        # the process's concurrent-work streams come before the context's own streams
        # (their hardware queues; maple_amd/streams.py)
        if torch.cuda.is_available():
            _streams.reserve(self.device)
        flat = (C.c_int * (3 * self.ncells))(*[v for d in dl for v in d])
        h = C.c_void_p()
        L.check(self.lib.maple_ctx_create(self.device.index or 0, self.ncells, flat, self.nlevels, prec, C.byref(h)))
        self._h = h
        self.sizes = [int(np.prod(d)) for d in dl]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.nelem = int(self.offsets[-1])
        self.generation = {"elastic": 0, "thermal": 0}
        self.last_iters = None
        self.last_status = 0
        self.set_smoother(smoother)
        if omega is None:
            omega = 1.6 if smoother == "l1-jacobi" else 0.6
        self.set_solver(omega, nsmooth, warm_start)
        if block_cg:
            self.set_cg(block=True)
        if islands:
            self.set_islands(True)

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self.lib.maple_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers ----------------------------------------------------------
    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _vec(self, t: torch.Tensor, n: int, name: str) -> torch.Tensor:
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
        if t.device != self.device:
            raise ValueError(f"{name} is on {t.device}, the batch is on {self.device}")
        if t.numel() != n:
            raise ValueError(f"{name} has {t.numel()} values, expected {n}")
        return t.reshape(-1).to(torch.float64).contiguous()

    def _empty(self, *shape):
        return torch.empty(*shape, dtype=torch.float64, device=self.device)

    # -- solver ------------------------------------------------------------
    def set_solver(self, omega: float = 1.6, nsmooth: int = 2, warm_start: bool = False):
        """Smoother weight, sweeps per V-cycle side, and start from the previous call's fields."""
        L.check(self.lib.maple_set_solver(self._h, float(omega), int(nsmooth), int(bool(warm_start))))

    def set_smoother(self, kind: str = "l1-jacobi"):
        """'jacobi' (weighted point Jacobi) or 'l1-jacobi' (absolute row sums; default)."""
        k = {"jacobi": L.MAPLE_SMOOTH_JACOBI, "l1-jacobi": L.MAPLE_SMOOTH_L1_JACOBI}[kind]
        L.check(self.lib.maple_set_smoother(self._h, k))

    def set_cg(self, flexible: Optional[bool] = None, block: bool = False):
        """Flexible (Polak-Ribiere) CG on/off (None = only in the mixed form), or block CG
        over the load cases of each cell."""
        mode = 2 if block else (-1 if flexible is None else int(bool(flexible)))
        L.check(self.lib.maple_set_cg(self._h, mode))

    def set_concurrency(self, on: bool = True):
        """Solve elastic and thermal at the same time (second stream and host thread) or in turn."""
        L.check(self.lib.maple_set_concurrency(self._h, int(bool(on))))

    def set_graphs(self, on: bool = True):
        """Run the CG iterations as captured hipGraphs (bitwise the same results)."""
        L.check(self.lib.maple_set_graphs(self._h, int(bool(on))))

    def graph_counts(self):
        """(segments captured, graph launches) so far."""
        a, b = C.c_longlong(0), C.c_longlong(0)
        L.check(self.lib.maple_graph_counts(self._h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def set_sweeps(self, pre: int = 2, post: int = 2):
        """Jacobi sweeps before / after the coarse correction (unequal: use flexible CG)."""
        L.check(self.lib.maple_set_sweeps(self._h, int(pre), int(post)))

    def set_chebyshev(self, lo: float = 0.0):
        """Smoother sweep weights from the Chebyshev nodes of [lo, 1] (0: every sweep omega)."""
        L.check(self.lib.maple_set_chebyshev(self._h, float(lo)))

    def set_islands(self, on: bool = True, threshold: float = 1e-4):
        """Island correction of the preconditioner: small solid islands (rho > threshold,
        at most 16 elements and 32 nodes) are solved exactly beside the V-cycle."""
        L.check(self.lib.maple_set_islands(self._h, int(bool(on)), float(threshold)))

    def num_islands(self) -> int:
        """Islands corrected in the last homogenization."""
        return int(self.lib.maple_num_islands(self._h))

    def set_mixed_fallback(self, iters: int = 300):
        """Mixed form: finish a solve on the FP64 V-cycle once it has run `iters` CG
        iterations (0 = never)."""
        L.check(self.lib.maple_set_mixed_fallback(self._h, int(iters)))

    def set_variant(self, kernel: int, variant: int):
        L.check(self.lib.maple_set_variant(self._h, int(kernel), int(variant)))

    def set_coarse_refresh(self, every: int = 1):
        """Rebuild the coarsest dense inverse every `every` homogenizations (1 = always)."""
        L.check(self.lib.maple_set_coarse_refresh(self._h, int(every)))

    def set_start_history(self, m: int = 0):
        """Start each solve from the energy-optimal combination of the last `m` solutions
        and a supplied start (0: off); clears the history."""
        L.check(self.lib.maple_set_start_history(self._h, int(m)))

    def set_element_matrices(self, cell: int, Ke24=None, Fe24=None, X0_24=None, Ke8=None, Fe8=None, X0_8=None):
        def p(a):
            if a is None:
                return None
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
            return a
        arrs = [p(a) for a in (Ke24, Fe24, X0_24, Ke8, Fe8, X0_8)]
        ptrs = [L.dptr(a) if a is not None else None for a in arrs]
        L.check(self.lib.maple_set_element_matrices(self._h, int(cell), *ptrs))

    def homogenize(self, rho: torch.Tensor, physics: Iterable[str] = ("elastic", "thermal"),
                   rtol: float = 1e-6, maxit: int = 500, strict: bool = True):
        """Returns (C_H (ncells,6,6) or None, kappa (ncells,3,3) or None, iterations (ncells,9))."""
        physics = tuple(physics)
        rho = self._vec(rho, self.nelem, "rho")
        mask = 0
        for p in physics:
            mask |= PHYS[p]
        CH = self._empty(self.ncells, 6, 6) if "elastic" in physics else None
        kap = self._empty(self.ncells, 3, 3) if "thermal" in physics else None
        iters = np.zeros(self.ncells * 9, dtype=np.int32)
        rc = self.lib.maple_homogenize(
            self._h, C.c_void_p(rho.data_ptr()), mask, float(rtol), int(maxit),
            C.c_void_p(CH.data_ptr()) if CH is not None else None,
            C.c_void_p(kap.data_ptr()) if kap is not None else None,
            iters.ctypes.data_as(L._IP), self._stream())
        self.last_iters = iters.reshape(self.ncells, 9)
        for p in physics:
            self.generation[p] += 1
        self.last_status = rc
        if rc != L.MAPLE_OK and (strict or rc not in (L.MAPLE_ENOCONV, L.MAPLE_EBREAKDOWN)):
            L.check(rc)
        return CH, kap, self.last_iters

    def sensitivity(self, G6: Optional[torch.Tensor] = None, G3: Optional[torch.Tensor] = None) -> torch.Tensor:
        """dJ/drho_e = <G6, S_e> + <G3, T_e> from the last homogenize (per-cell G)."""
        g6 = self._vec(G6, 36 * self.ncells, "G6") if G6 is not None else None
        g3 = self._vec(G3, 9 * self.ncells, "G3") if G3 is not None else None
        out = self._empty(self.nelem)
        L.check(self.lib.maple_sensitivity(
            self._h, C.c_void_p(g6.data_ptr()) if g6 is not None else None,
            C.c_void_p(g3.data_ptr()) if g3 is not None else None, C.c_void_p(out.data_ptr()), self._stream()))
        return out

    def filter(self, v: torch.Tensor, k: int = 3) -> torch.Tensor:
        v = self._vec(v, self.nelem, "v")
        out = self._empty(self.nelem)
        L.check(self.lib.maple_filter_box(self._h, C.c_void_p(v.data_ptr()), C.c_void_p(out.data_ptr()), int(k),
                                          self._stream()))
        return out

    def oc_update(self, x: torch.Tensor, dJdx: torch.Tensor, volfrac, move=0.2, xmin=1e-3, nbisect: int = 60):
        x = self._vec(x, self.nelem, "x")
        dJ = self._vec(dJdx, self.nelem, "dJdx")

        def per_cell(v):
            a = np.broadcast_to(np.asarray(v, dtype=np.float64), (self.ncells,)).copy()
            return a
        vf, mv, xm = per_cell(volfrac), per_cell(move), per_cell(xmin)
        out = self._empty(self.nelem)
        L.check(self.lib.maple_oc_update(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(dJ.data_ptr()), L.dptr(vf),
                                         L.dptr(mv), L.dptr(xm), int(nbisect), C.c_void_p(out.data_ptr()),
                                         self._stream()))
        return out

    def symmetrize(self, v: torch.Tensor, kind: str = "none", axes: Sequence[int] = ()) -> torch.Tensor:
        """Mean over reflections about the centre planes of `axes` ("mirror") or the 48
        cube symmetries ("cubic") of every cell."""
        v = self._vec(v, self.nelem, "v")
        out = torch.empty_like(v)
        k = {"none": 0, "mirror": 1, "cubic": 2}[kind]
        mask = sum(1 << int(a) for a in axes)
        L.check(self.lib.maple_symmetrize(self._h, k, mask, C.c_void_p(v.data_ptr()), C.c_void_p(out.data_ptr()),
                                          self._stream()))
        return out

    def rhs(self, physics: str, rho: torch.Tensor) -> torch.Tensor:
        rho = self._vec(rho, self.nelem, "rho")
        F = self._empty(NCASE[physics], NDOF[physics], self.nelem)
        L.check(self.lib.maple_rhs(self._h, PHYS[physics], C.c_void_p(rho.data_ptr()), C.c_void_p(F.data_ptr()),
                                   self._stream()))
        return F

    def apply_operator(self, physics: str, rho: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
        """v = K(rho) u for u of shape (ncase, ndof, Etot), ncase <= the physics' case count."""
        rho = self._vec(rho, self.nelem, "rho")
        nd = NDOF[physics]
        if u.dim() != 3 or u.shape[1] != nd or u.shape[2] != self.nelem:
            raise ValueError(f"u must have shape (ncase, {nd}, {self.nelem})")
        u = u.to(torch.float64).contiguous()
        v = torch.empty_like(u)
        L.check(self.lib.maple_apply_operator(self._h, PHYS[physics], C.c_void_p(rho.data_ptr()), int(u.shape[0]),
                                              C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()), self._stream()))
        return v

    def fields(self, physics: str) -> torch.Tensor:
        u = self._empty(NCASE[physics], NDOF[physics], self.nelem)
        L.check(self.lib.maple_get_fields(self._h, PHYS[physics], C.c_void_p(u.data_ptr()), self._stream()))
        return u

    def set_fields(self, physics: str, u: torch.Tensor):
        u = self._vec(u, NCASE[physics] * NDOF[physics] * self.nelem, "u")
        L.check(self.lib.maple_set_fields(self._h, PHYS[physics], C.c_void_p(u.data_ptr()), self._stream()))

    def set_start(self, physics: str, u: torch.Tensor):
        """Candidate start of the next solve (ncase, ndof, Etot); per (cell, case) the solve
        starts from the best of zero, the previous fields (warm start) and this."""
        u = self._vec(u, NCASE[physics] * NDOF[physics] * self.nelem, "u")
        L.check(self.lib.maple_set_start(self._h, PHYS[physics], C.c_void_p(u.data_ptr()), self._stream()))

    def start_sources(self, physics: str) -> np.ndarray:
        """Start chosen by the last solve, (ncells, ncase): 0 zero, 1 previous fields, 2 supplied,
        3 projected on earlier solutions."""
        out = (C.c_int * (self.ncells * NCASE[physics]))()
        L.check(self.lib.maple_start_sources(self._h, PHYS[physics], out))
        return np.array(out, dtype=np.int32).reshape(self.ncells, NCASE[physics])

    def num_nodes(self, level: int) -> int:
        return int(self.lib.maple_num_nodes(self._h, int(level)))

    def prolong(self, physics: str, level: int, coarse: torch.Tensor) -> torch.Tensor:
        """Periodic trilinear prolongation of (ncase, ndof, N_level) fields to level - 1."""
        nc, nd = NCASE[physics], NDOF[physics]
        coarse = self._vec(coarse, nc * nd * self.num_nodes(level), "coarse")
        fine = self._empty(nc, nd, self.num_nodes(level - 1))
        L.check(self.lib.maple_prolong(self._h, PHYS[physics], int(level), C.c_void_p(coarse.data_ptr()),
                                       C.c_void_p(fine.data_ptr()), self._stream()))
        return fine

    # -- additive-manufacturing penalties ---------------------------------------
    def penalty_overhang(self, rho: torch.Tensor, overhang_deg: float = 45.0, periodic_xy: bool = True,
                         grad: bool = True):
        """printability_penalty_overhang of every cell: (P (ncells,), dP/drho (Etot,) or None)."""
        rho = self._vec(rho, self.nelem, "rho")
        P = self._empty(self.ncells)
        dP = self._empty(self.nelem) if grad else None
        L.check(self.lib.maple_penalty_overhang(self._h, C.c_void_p(rho.data_ptr()), float(overhang_deg),
                                                int(bool(periodic_xy)), C.c_void_p(P.data_ptr()),
                                                C.c_void_p(dP.data_ptr()) if dP is not None else None,
                                                self._stream()))
        return P, dP

    def set_cavity(self, void_thresh: float = 0.5, beta: float = 20.0, min_channel_rad: int = 0, rad: int = 1,
                   iters: int = 0, periodic_xyz: bool = True, seed_faces: str = "all", tiles=(1, 1, 1),
                   focus_center: bool = True):
        """Configure (and allocate) the trapped-void penalty: tiles all 1 is
        cavity_penalty_closed_void, otherwise cavity_penalty_periodic_array on the tiled cell."""
        t = (int(tiles),) * 3 if isinstance(tiles, (int, np.integer)) else tuple(int(v) for v in tiles)
        o = L.CavityOpts(float(void_thresh), float(beta), int(min_channel_rad), int(rad), int(iters),
                         int(bool(periodic_xyz)), L.SEED_FACES[str(seed_faces).lower()], (C.c_int * 3)(*t),
                         int(bool(focus_center)))
        L.check(self.lib.maple_set_cavity(self._h, C.byref(o)))

    def penalty_cavity(self, rho: torch.Tensor, grad: bool = True):
        """The configured trapped-void penalty: (P (ncells,), dP/drho (Etot,) or None)."""
        rho = self._vec(rho, self.nelem, "rho")
        P = self._empty(self.ncells)
        dP = self._empty(self.nelem) if grad else None
        L.check(self.lib.maple_penalty_cavity(self._h, C.c_void_p(rho.data_ptr()), C.c_void_p(P.data_ptr()),
                                              C.c_void_p(dP.data_ptr()) if dP is not None else None, self._stream()))
        return P, dP

    # -- multigrid internals -------------------------------------------------
    def _prec(self, precision: str) -> int:
        """Hierarchy of the multigrid internals: "fp64", or the mixed form's FP32 one
        ("mixed", also spelled "fp32")."""
        return {"fp64": L.MAPLE_PREC_FP64, "mixed": L.MAPLE_PREC_MIXED, "fp32": L.MAPLE_PREC_MIXED}[precision]

    def mg_setup(self, physics: str, rho: torch.Tensor, precision: str = "fp64"):
        """Build the hierarchy of rho without solving (the FP64 level 0 keeps reading
        rho, so the returned tensor must stay alive while level 0 is used)."""
        rho = self._vec(rho, self.nelem, "rho")
        L.check(self.lib.maple_mg_setup(self._h, PHYS[physics], C.c_void_p(rho.data_ptr()), self._prec(precision),
                                        self._stream()))
        self._mg_rho = rho
        return rho

    def mg_stencil(self, physics: str, level: int, precision: str = "fp64") -> torch.Tensor:
        """(27, nd, nd, N_level): entry [d, i, j, n] couples node n (row i) to node
        n + off(d) (column j), d = (dx+1)*9 + (dy+1)*3 + (dz+1)."""
        nd = NDOF[physics]
        out = self._empty(27, nd, nd, self.num_nodes(level))
        L.check(self.lib.maple_mg_stencil(self._h, PHYS[physics], self._prec(precision), int(level),
                                          C.c_void_p(out.data_ptr()), self._stream()))
        return out

    def mg_dinv(self, physics: str, level: int, precision: str = "fp64") -> torch.Tensor:
        """Inverse smoother diagonal (nd, N_level)."""
        out = self._empty(NDOF[physics], self.num_nodes(level))
        L.check(self.lib.maple_mg_dinv(self._h, PHYS[physics], self._prec(precision), int(level),
                                       C.c_void_p(out.data_ptr()), self._stream()))
        return out

    def mg_apply(self, physics: str, level: int, u: torch.Tensor, f: Optional[torch.Tensor] = None, mode: int = 0,
                 precision: str = "fp64") -> torch.Tensor:
        """mode 0: A u; 2: f - A u; 3: one smoother sweep u + omega Dinv (f - A u); fields
        of shape (ncase, nd, N_level)."""
        n = NCASE[physics] * NDOF[physics] * self.num_nodes(level)
        u = self._vec(u, n, "u")
        fv = self._vec(f, n, "f") if f is not None else None
        out = self._empty(NCASE[physics], NDOF[physics], self.num_nodes(level))
        L.check(self.lib.maple_mg_apply(self._h, PHYS[physics], self._prec(precision), int(level), int(mode),
                                        C.c_void_p(u.data_ptr()), C.c_void_p(fv.data_ptr()) if fv is not None else None,
                                        C.c_void_p(out.data_ptr()), self._stream()))
        return out

    def mg_restrict(self, physics: str, level: int, fine: torch.Tensor) -> torch.Tensor:
        """P^T fine from level - 1 to level, (ncase, nd, N_level)."""
        fine = self._vec(fine, NCASE[physics] * NDOF[physics] * self.num_nodes(level - 1), "fine")
        out = self._empty(NCASE[physics], NDOF[physics], self.num_nodes(level))
        L.check(self.lib.maple_mg_restrict(self._h, PHYS[physics], int(level), C.c_void_p(fine.data_ptr()),
                                           C.c_void_p(out.data_ptr()), self._stream()))
        return out

    def mg_vcycle(self, physics: str, r: torch.Tensor, precision: str = "fp64") -> torch.Tensor:
        """One V-cycle z = M r at level 0 (every load case)."""
        r = self._vec(r, NCASE[physics] * NDOF[physics] * self.nelem, "r")
        out = self._empty(NCASE[physics], NDOF[physics], self.nelem)
        L.check(self.lib.maple_mg_vcycle(self._h, PHYS[physics], self._prec(precision), C.c_void_p(r.data_ptr()),
                                         C.c_void_p(out.data_ptr()), self._stream()))
        return out

    def mg_coarsest_inverse(self, physics: str, cell: int = 0) -> torch.Tensor:
        n = int(self.lib.maple_mg_coarsest_order(self._h, PHYS[physics], int(cell)))
        out = self._empty(n, n)
        L.check(self.lib.maple_mg_coarsest_inverse(self._h, PHYS[physics], int(cell), C.c_void_p(out.data_ptr()),
                                                   self._stream()))
        return out

    # -- timing -------------------------------------------------------------
    def timing(self, enable: bool):
        L.check(self.lib.maple_timing_enable(self._h, int(bool(enable))))

    def timing_read(self, kind: int = 0, work: bool = False):
        """(launches, total ms) of the timed level-0 operator launches since the last
        read; with work=True also the elements x load cases they processed."""
        cnt = C.c_longlong(0)
        ms = C.c_double(0.0)
        wk = C.c_longlong(0)
        L.check(self.lib.maple_timing_read(self._h, int(kind), C.byref(cnt), C.byref(ms), C.byref(wk)))
        if work:
            return int(cnt.value), float(ms.value), int(wk.value)
        return int(cnt.value), float(ms.value)
