"""Objective grammar and terms (the reference's objective.py:11-251), plus bulk modulus.

Two forms of the same functions:

* torch versions (``stiffness_objective`` ...) with the reference's names,
  argument meaning and errors, differentiable with autograd; they run on
  whatever device C_H lives on.
* ``evaluate_np``: closed-form values and gradients dJ/dC_H, dJ/dkappa in
  float64 numpy on 6x6 / 3x3 host arrays, used by the GPU OC driver to avoid
  a round of tiny device kernels per iteration.  With A = sym(C_H) + 1e-9 I
  and S = A^-1, d(1/S_ii)/dA = S[:,i] S[i,:] / S_ii^2 (from dS = -S dA S).

Terms: stiffness = mean_{i<3} 1/S_ii, shear = mean_{i=3..5} 1/S_ii,
thermal = trace(kappa) (objective.py:11-41); bulk = (1/9) sum_{i,j<3} C_H[i,j]
(Voigt bulk modulus; an extension the reference does not have).
"""
from __future__ import annotations

import re
from typing import Dict, Mapping, Optional, Tuple

import numpy as np
import torch

# Accepted names of each term (objective.py:65-82 lists the reference's; "bulk" and its
# names are this build's extension).
_NAMES = {
    "stiffness": ("stiffness", "stiff", "e", "young", "youngs"),
    "shear": ("shear", "g"),
    "thermal": ("thermal", "kappa", "conductivity", "heat"),
    "bulk": ("bulk", "k", "bulk_modulus"),
}
_TERM_OF = {alias: term for term, names in _NAMES.items() for alias in names}

# One weighted term: "name", "name:w", "name=w", "w*name" or "name*w".
_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
_WORD = r"[^\s:=*,+]+"
_TERM_RE = re.compile(rf"^(?:(?P<w1>{_NUM})\*(?P<n1>{_WORD})|(?P<n2>{_WORD})\*(?P<w2>{_NUM})|"
                      rf"(?P<n3>{_WORD})(?:[:=](?P<w3>{_NUM}))?)$")


def _term(name: str) -> str:
    key = str(name).strip().lower()
    if key not in _TERM_OF:
        raise ValueError(f"objective term {name!r} is not one of {sorted(_NAMES)} (or their aliases)")
    return _TERM_OF[key]


def _is_number(text: str) -> bool:
    return re.fullmatch(_NUM, text) is not None


def parse_objective_spec(spec: Optional[str]) -> Dict[str, float]:
    """Weights of an objective spec (objective.py:93-157): a single term name, or terms
    joined by ',' or '+', each "name", "name:w", "name=w", "w*name" or "name*w" (spaces
    ignored; a term named twice adds up; weights are used as given).  An empty spec is
    the reference's default, shear."""
    text = "" if spec is None else str(spec).strip()
    if not text:
        return {"shear": 1.0}
    if text.lower() in _TERM_OF:
        return {_TERM_OF[text.lower()]: 1.0}
    weights: Dict[str, float] = {}
    pieces = [p for p in re.split(r"[,+]", text.replace(" ", "")) if p]
    if not pieces:
        raise ValueError(f"objective spec {spec!r} names no term")
    for piece in pieces:
        m = _TERM_RE.match(piece)
        name = m and (m["n1"] or m["n2"] or m["n3"])
        if not m or _is_number(name):
            raise ValueError(f"cannot read {piece!r} of objective spec {spec!r}: write a term as "
                             f"'name', 'name:w', 'name=w' or 'w*name'")
        w = m["w1"] or m["w2"] or m["w3"]
        term = _term(name)
        weights[term] = weights.get(term, 0.0) + (float(w) if w is not None else 1.0)
    return weights


def objective_requires(spec: str) -> Tuple[bool, bool]:
    terms = parse_objective_spec(spec)
    return (any(k in ("stiffness", "shear", "bulk") for k in terms), "thermal" in terms)


def _compliance(CH: torch.Tensor) -> torch.Tensor:
    C = 0.5 * (CH + CH.T)
    return torch.linalg.inv(C + 1e-9 * torch.eye(6, device=C.device, dtype=C.dtype))


def stiffness_objective(CH: torch.Tensor) -> torch.Tensor:
    S = _compliance(CH)
    return (1.0 / S[0, 0] + 1.0 / S[1, 1] + 1.0 / S[2, 2]) / 3.0


def shear_objective(CH: torch.Tensor) -> torch.Tensor:
    S = _compliance(CH)
    return (1.0 / S[3, 3] + 1.0 / S[4, 4] + 1.0 / S[5, 5]) / 3.0


def bulk_objective(CH: torch.Tensor) -> torch.Tensor:
    return CH[:3, :3].sum() / 9.0


def thermal_objective(kappa: torch.Tensor) -> torch.Tensor:
    if kappa.dim() < 2 or kappa.shape[-1] != kappa.shape[-2]:
        raise ValueError(f"kappa must be square, got {tuple(kappa.shape)}")
    return torch.diagonal(kappa, dim1=-2, dim2=-1).sum()


def objective_from_spec(*, obj: str, CH: Optional[torch.Tensor] = None, kappa: Optional[torch.Tensor] = None,
                        normalize_by: Optional[Mapping[str, float]] = None, return_parts: bool = False):
    """Weighted multi-objective J (objective.py:180-251)."""
    terms = parse_objective_spec(obj)
    raw_parts, used_parts = {}, {}
    J = None
    for term, w in terms.items():
        if term in ("stiffness", "shear", "bulk"):
            if CH is None:
                raise ValueError(f"Objective '{obj}' requires CH (mechanical), but CH=None")
            raw = {"stiffness": stiffness_objective, "shear": shear_objective, "bulk": bulk_objective}[term](CH)
        else:
            if kappa is None:
                raise ValueError(f"Objective '{obj}' requires kappa (thermal), but kappa=None")
            raw = thermal_objective(kappa)
        raw_parts[term] = raw
        used = raw / max(abs(float(normalize_by[term])), 1e-12) if normalize_by and term in normalize_by else raw
        used_parts[term] = used
        J = float(w) * used if J is None else J + float(w) * used
    if J is None:
        raise ValueError(f"Empty objective spec: {obj!r}")
    return (J, raw_parts, used_parts) if return_parts else J


# ---- closed form in numpy (host) ---------------------------------------------

def _inv_diag_np(CH, idx):
    A = 0.5 * (CH + CH.T) + 1e-9 * np.eye(6)
    S = np.linalg.inv(A)
    val = float(np.mean([1.0 / S[i, i] for i in idx]))
    G = sum(np.outer(S[:, i], S[i, :]) / S[i, i] ** 2 for i in idx) / len(idx)
    return val, 0.5 * (G + G.T)


def evaluate_np(weights: Mapping[str, float], CH: Optional[np.ndarray] = None, kappa: Optional[np.ndarray] = None):
    """(J, dJ/dC_H (6,6), dJ/dkappa (3,3)) for one cell, float64."""
    J = 0.0
    G6 = np.zeros((6, 6))
    G3 = np.zeros((3, 3))
    for term, w in weights.items():
        if term == "thermal":
            J += w * float(np.trace(kappa))
            G3 += w * np.eye(3)
        elif term == "bulk":
            J += w * float(np.sum(CH[:3, :3]) / 9.0)
            G6[:3, :3] += w / 9.0
        else:
            v, g = _inv_diag_np(CH, (0, 1, 2) if term == "stiffness" else (3, 4, 5))
            J += w * v
            G6 += w * g
    return J, G6, G3
