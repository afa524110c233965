"""GeoPhyInv.jl's staggered-grid FDTD engine on MI355X: host API over libgpi.

The names mirror the reference's Julia API: ``SeisForwExpt``, ``update`` (``update!``),
``AcousticMedium``/``ElasticMedium``, ``AGeomss``/``AGeom``, ``Srcs``/``Recs``
(``SrcWav``/``Records``), ``ricker``, ``FdtdAcoustic``/``FdtdElastic``.
"""
from . import database as _database
from .born import Born, LinearMap
from . import func_grad as _func_grad
from .database import AGeomss, L2DistLoss, Records, Recs, SrcWav, Srcs
from .fdtd import FdtdAcoustic, FdtdElastic, SeisForwExpt, update
from .fields import FIELD_ID, FIELDS, field_shape, get_mgrid
from .fwi import SeisInvExpt
from .grid import Grid
from .media import AcousticMedium, ElasticMedium, Medium, padarray
from .chainrules import rrule
from .wavelets import findfreq, ricker

AGeom = list


def lossvalue(*args, **kw):
    """``lossvalue(loss, dobs, dcal)`` (database.jl:357-368) or
    ``lossvalue(m, loss, dobs, pa[, mparams])`` (func_grad.jl:1-9)."""
    if len(args) >= 4:
        return _func_grad.lossvalue(*args, **kw)
    return _database.lossvalue(*args, **kw)


def gradient(*args, **kw):
    """``gradient!(buffer, loss, dobs, dcal)`` (database.jl:371-380) or
    ``gradient!(g, m, loss, dobs, pa[, mparams])`` (func_grad.jl:11-49)."""
    if len(args) >= 5:
        return _func_grad.gradient(*args, **kw)
    return _database.gradient(*args, **kw)


__all__ = [
    "AGeom", "AGeomss", "AcousticMedium", "Born", "ElasticMedium", "LinearMap", "FIELDS", "FIELD_ID", "FdtdAcoustic",
    "FdtdElastic", "Grid", "L2DistLoss", "Medium", "Records", "Recs", "SeisForwExpt", "SeisInvExpt", "SrcWav", "Srcs",
    "field_shape", "findfreq", "get_mgrid", "gradient", "lossvalue", "padarray", "ricker", "rrule", "update",
]
