"""gfem: GmshFEM's Formulation::assemble() hot path on MI355X (gfx950).

Python face of the C++ host API (fem-g_amd/host) and of the C ABI
(include/gfem.h, libgfem.so).  The native libraries are built in-tree by
fem-g_amd/build.py; importing this package without them raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _gfem_host as _h  # noqa: F401
except ImportError as e:  # pragma: no cover - exercised only on a broken build
    raise ImportError(
        "gfem native extension not built: run `python fem-g_amd/build.py` "
        f"(or __graft_entry__.build()) first ({e})") from e

from ._gfem_host import (Domain, Field, ManyInstance, Model, ScalarFunction, Server, Unknown, abs,  # noqa: F401,E402
                         assemble_many, cos, dof, dt2_dof, dt_dof, exp, from_arrays, gauss_rule, grad,
                         hier_basis, hier_keys, hilbert_index, hilbert_key, lattice_nodes, pow, read_msh,
                         set_current_model, set_device, device_context_handle,
                         sin, sqrt, tf, unit_cube, unit_square, write_msh, x, y, z)

_FORMULATIONS = {
    "f64": _h.Formulation_f64,
    "f32": _h.Formulation_f32,
    "c128": _h.Formulation_c128,
    "c64": _h.Formulation_c64,
}

LIBGFEM = os.path.join(_HERE, "libgfem.so")


def Formulation(name, scalar="f64"):
    """problem::Formulation<T> with T named by `scalar` (f64, f32, c128, c64)."""
    try:
        cls = _FORMULATIONS[scalar]
    except KeyError:
        raise ValueError(f"scalar must be one of {sorted(_FORMULATIONS)}") from None
    return cls(name)


def load_c_abi():
    """ctypes handle on libgfem.so (the C ABI declared in include/gfem.h)."""
    return ctypes.CDLL(LIBGFEM)


def c_abi_symbols(header=None):
    """Entry points declared in include/gfem.h."""
    import re
    if header is None:
        header = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "gfem.h")
    text = open(header).read()
    return sorted(set(re.findall(r"\b(gfem_[a-z_0-9]+)\s*\(", text)))
