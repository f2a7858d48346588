"""Settings of an OpenFOAM case for the steady SIMPLE path (host side).

What the reference reads from a case directory before it runs the steady
incompressible loop, restricted to this path:

* ``constant/polyMesh``: the mesh, processed as ``cfdProcessOpenFoamMesh``
  does (``mesh.process_mesh``);
* ``0/U`` and ``0/p``: the uniform internal fields (``cfdDefineEquation.m:68-82``)
  and, per patch, the boundary condition type and value
  (``cfdReadTimeDirectory.m:13-133``);
* ``constant/transportProperties``: ``nu`` (tutorials written
  ``div(U*U) = laplacian(nu*U) ...``, density 1), or ``mu`` and ``rho``
  (tutorials written ``div(rho*U*U) = laplacian(mu*U) ...``: mu is the
  diffusion coefficient, rho the density of the convected mass flux);
* ``system/fvSolution``: ``nCorrectors`` and ``pRefCell`` of the SIMPLE block
  (``cfdReadSystem.m:150-190``) and the relaxation factors, p from ``fields``
  and U from ``equations`` (``cfdDefineEquation.m:84-99``, 0.7 when absent).

The pressure reference cell is only used when every patch is a wall, empty or
symmetry patch (``cfdReadSystem.m:238-251``); otherwise there is none.  A
``pRefCell`` of 0 becomes cell 1, any other value is kept as a 1-based index
(``cfdReadSystem.m:170-178``), i.e. 0-based cell v - 1.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

from .fieldio import read_field
from .mesh import Mesh, process_mesh
from .polymesh import read_polymesh

PLANE = ("empty", "symmetry")


@dataclass
class FoamCase:
    mesh: Mesh
    bcs: dict                     # patch -> (U type, U value (3,), p type, p value)
    nu: float                     # the laplacian coefficient (nu, or mu)
    u_init: tuple
    p_init: float
    n_correctors: int = 1
    urf_u: float = 0.7
    urf_p: float = 0.7
    p_ref: int = -1               # 0-based, -1 none
    rho: float = 1.0
    extra: dict = field(default_factory=dict)

    def simple_kwargs(self) -> dict:
        """keyword arguments of SimpleCase (GPU) and oracle.SimpleCase"""
        return dict(nu=self.nu, urf_u=self.urf_u, urf_p=self.urf_p,
                    n_correctors=self.n_correctors, p_ref=self.p_ref, u_init=self.u_init,
                    p_init=self.p_init, rho=self.rho)


def _strip(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _block(text: str, name: str) -> str | None:
    """body of the first `name { ... }` dictionary in text"""
    m = re.search(r"(?<![\w.])" + re.escape(name) + r"\s*\{", text)
    if not m:
        return None
    depth, i = 1, m.end()
    while depth:
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
        i += 1
    return text[m.end():i - 1]


def _entry(body: str | None, key: str) -> str | None:
    if body is None:
        return None
    m = re.search(r"(?<![\w.])" + re.escape(key) + r"\s+([^;{}]*);", body)
    return m.group(1).strip() if m else None


def read_case(case_dir: str, apex: str = "reference") -> FoamCase:
    """The steady SIMPLE settings of the OpenFOAM case in case_dir."""
    mesh = process_mesh(read_polymesh(case_dir), apex=apex)
    U0 = read_field(os.path.join(case_dir, "0", "U"))
    p0 = read_field(os.path.join(case_dir, "0", "p"))
    bcs = {}
    for pt in mesh.patches:
        ue, pe = U0.boundary[pt.name], p0.boundary[pt.name]
        uval = tuple(float(v) for v in ue.value) if ue.value is not None else (0.0, 0.0, 0.0)
        pval = float(pe.value) if pe.value is not None else 0.0
        bcs[pt.name] = (ue.type, uval, pe.type, pval)
    tp = _strip(open(os.path.join(case_dir, "constant", "transportProperties")).read())

    def prop(name):
        mm = re.search(r"(?<![\w.])" + name + r"\b[^;]*?([-+0-9.eE]+)\s*;", tp)
        return float(mm.group(1)) if mm else None

    nu, rho = prop("nu"), 1.0
    if nu is None:
        nu, rho = prop("mu"), prop("rho")
    fvs = _strip(open(os.path.join(case_dir, "system", "fvSolution")).read())
    simple = _block(fvs, "SIMPLE")
    nc = _entry(simple, "nCorrectors")
    rf = _block(fvs, "relaxationFactors")
    urf_p = _entry(_block(rf, "fields") if rf else None, "p")
    urf_u = _entry(_block(rf, "equations") if rf else None, "U")
    p_ref = -1
    if all(pt.type in ("wall",) + PLANE for pt in mesh.patches):
        v = _entry(simple, "pRefCell")
        if v is not None:
            v1 = int(float(v))
            p_ref = (v1 + 1 if v1 == 0 else v1) - 1
    u_init = tuple(float(v) for v in U0.internal)
    return FoamCase(mesh=mesh, bcs=bcs, nu=nu, u_init=u_init, p_init=float(p0.internal),
                    n_correctors=int(float(nc)) if nc is not None else 1,
                    urf_u=float(urf_u) if urf_u is not None else 0.7,
                    urf_p=float(urf_p) if urf_p is not None else 0.7, p_ref=p_ref, rho=rho)
