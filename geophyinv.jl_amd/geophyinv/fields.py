"""Field names, ids and staggered grids (src/fields.jl).

Field ids follow ``enum gpi_field`` of include/gpi.h.  ``get_mgrid`` restates
src/fields.jl:92-671: every field lives on a grid derived from the extended stress grid
by one of four per-axis rules (h = order - 1):

  N  the stress grid itself                       (n points)
  A  shifted by -h/2 cells, h points longer        (normal velocities)
  H  shifted by +h/2 cells, h points shorter       (half-staggered derivatives, shear stress)
  I  shifted by +h cells,   2h points shorter      (interior of a centred derivative)
"""
from __future__ import annotations

from .grid import Grid

FIELDS = [
    "p", "vx", "vy", "vz", "tauxx", "tauyy", "tauzz", "tauxy", "tauxz", "tauyz",
    "dpdx", "dpdy", "dpdz",
    "dvxdx", "dvydy", "dvzdz", "dvxdy", "dvxdz", "dvydx", "dvydz", "dvzdx", "dvzdy",
    "dtauxxdx", "dtauyydy", "dtauzzdz", "dtauxydx", "dtauxydy", "dtauxzdx", "dtauxzdz",
    "dtauyzdy", "dtauyzdz",
]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}

COEFS = [
    "dtinvavxirho", "dtinvavyirho", "dtinvavzirho", "dtK", "dtM", "dtlambda", "dtavmu",
    "dtavxzimu", "dtavxyimu", "dtavyzimu", "rho", "invK", "invlambda", "invmu",
]
COEF_ID = {name: i for i, name in enumerate(COEFS)}

# per-axis codes (z, y, x); '.' = absent y axis in 2-D
_AC2 = {"p": "N.N", "vx": "N.A", "vz": "A.N", "dpdx": "I.H", "dpdz": "H.I",
        "dvxdx": "N.N", "dvzdz": "N.N"}
_AC3 = {"p": "NNN", "vx": "NNA", "vy": "NAN", "vz": "ANN", "dpdx": "IIH", "dpdy": "IHI",
        "dpdz": "HII", "dvxdx": "NNN", "dvydy": "NNN", "dvzdz": "NNN"}
_EL2 = {"tauxx": "N.N", "tauzz": "N.N", "tauxz": "H.H", "vx": "N.A", "vz": "A.N",
        "dtauxxdx": "I.H", "dtauzzdz": "H.I", "dtauxzdx": "H.I", "dtauxzdz": "I.H",
        "dvxdx": "N.N", "dvzdz": "N.N", "dvxdz": "H.H", "dvzdx": "H.H"}
_EL3 = {"tauxx": "NNN", "tauyy": "NNN", "tauzz": "NNN", "tauxy": "IHH", "tauxz": "HIH",
        "tauyz": "HHI", "vx": "NNA", "vy": "NAN", "vz": "ANN",
        "dtauxxdx": "IIH", "dtauyydy": "IHI", "dtauzzdz": "HII", "dtauxydx": "IHI",
        "dtauxydy": "IIH", "dtauxzdx": "HII", "dtauxzdz": "IIH", "dtauyzdy": "HII",
        "dtauyzdz": "IHI", "dvxdx": "NNN", "dvydy": "NNN", "dvzdz": "NNN", "dvxdy": "IHH",
        "dvxdz": "HIH", "dvydx": "IHH", "dvydz": "HHI", "dvzdx": "HIH", "dvzdy": "HHI"}


def _table(physics: int, ndim: int) -> dict:
    if physics == 0:
        return _AC2 if ndim == 2 else _AC3
    return _EL2 if ndim == 2 else _EL3


def fields_of(physics: int, ndim: int, kind: str = "") -> list[str]:
    """``Fields(attrib_mod, s)`` (src/fields.jl:9-25): fields of a physics containing ``kind``,
    in the reference's (alphabetical) order."""
    return sorted(f for f in _table(physics, ndim) if kind in f)


def wavefields_of(physics: int, ndim: int) -> list[str]:
    return [f for f in fields_of(physics, ndim) if f not in fields_of(physics, ndim, "d")]


def derivative_fields(physics: int, ndim: int) -> list[str]:
    """Derivative fields that carry CPML memory (``get_pml``, src/fdtd/cpml.jl:109-120)."""
    return fields_of(physics, ndim, "d")


def has_field(physics: int, ndim: int, field: str) -> bool:
    return field in _table(physics, ndim)


def _codes(physics: int, ndim: int, field: str) -> str:
    t = _table(physics, ndim)
    if field not in t:
        raise KeyError(f"field {field!r} does not exist for this physics/ndim")
    return t[field]


def field_shape(physics: int, ndim: int, order: int, n, field: str) -> tuple:
    """Array size of a field on the extended grid ``n`` = (nz, [ny,] nx) (src/fields.jl:48-86)."""
    h = order - 1
    c = _codes(physics, ndim, field)
    axes = [0, 2] if ndim == 2 else [0, 1, 2]
    if len(n) != ndim:
        raise ValueError("n must have one size per axis")
    delta = {"N": 0, "A": h, "H": -h, "I": -2 * h}
    return tuple(int(nn) + delta[c[a]] for nn, a in zip(n, axes))


def get_mgrid(field: str, physics: int, grids: list[Grid], order: int) -> list[Grid]:
    """Staggered grid of ``field`` given the extended stress grid (z, [y,] x)."""
    h = order - 1
    ndim = len(grids)
    c = _codes(physics, ndim, field)
    axes = [0, 2] if ndim == 2 else [0, 1, 2]
    out = []
    for g, a in zip(grids, axes):
        code = c[a]
        if code == "N":
            out.append(g)
        elif code == "A":
            out.append(Grid(g.first - 0.5 * g.step * h, g.step, g.n + h))
        elif code == "H":
            out.append(Grid(g.first + 0.5 * g.step * h, g.step, g.n - h))
        else:
            out.append(Grid(g.first + g.step * h, g.step, g.n - 2 * h))
    return out


def deriv_axis(field: str) -> int:
    """Axis (0 z, 1 y, 2 x) along which a derivative field differentiates."""
    if not field.startswith("d"):
        raise ValueError(field)
    return {"z": 0, "y": 1, "x": 2}[field[-1]]


# grid of each coefficient array (src/fdtd/medium.jl:81-95)
def coef_grid(physics: int, ndim: int, coef: str) -> str | None:
    ac = physics == 0
    table = {
        "dtinvavxirho": "dpdx" if ac else "dtauxxdx",
        "dtinvavyirho": ("dpdy" if ac else "dtauyydy") if ndim == 3 else None,
        "dtinvavzirho": "dpdz" if ac else "dtauzzdz",
        "dtK": "p" if ac else None,
        "dtM": None if ac else "tauxx",
        "dtlambda": None if ac else "tauxx",
        "dtavmu": "dvxdz" if (not ac and ndim == 2) else None,
        "dtavxzimu": "dvxdz" if (not ac and ndim == 3) else None,
        "dtavxyimu": "dvxdy" if (not ac and ndim == 3) else None,
        "dtavyzimu": "dvydz" if (not ac and ndim == 3) else None,
        "rho": "p" if ac else "tauxx",
        "invK": "p" if ac else None,
        "invlambda": None if ac else "tauxx",
        "invmu": None if ac else "tauxx",
    }
    return table[coef]
