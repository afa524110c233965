"""GeoPhyInv.jl's staggered-grid FDTD engine on MI355X: host API over libgpi.

The names mirror the reference's Julia API: ``SeisForwExpt``, ``update`` (``update!``),
``AcousticMedium``/``ElasticMedium``, ``AGeomss``/``AGeom``, ``Srcs``/``Recs``
(``SrcWav``/``Records``), ``ricker``, ``FdtdAcoustic``/``FdtdElastic``.
"""
from .database import AGeomss, Records, Recs, SrcWav, Srcs, gradient, lossvalue
from .fdtd import FdtdAcoustic, FdtdElastic, SeisForwExpt, update
from .fields import FIELD_ID, FIELDS, field_shape, get_mgrid
from .grid import Grid
from .media import AcousticMedium, ElasticMedium, Medium, padarray
from .wavelets import findfreq, ricker

AGeom = list

__all__ = [
    "AGeom", "AGeomss", "AcousticMedium", "ElasticMedium", "FIELDS", "FIELD_ID", "FdtdAcoustic",
    "FdtdElastic", "Grid", "Medium", "Records", "Recs", "SeisForwExpt", "SrcWav", "Srcs",
    "field_shape", "findfreq", "get_mgrid", "gradient", "lossvalue", "padarray", "ricker", "update",
]
