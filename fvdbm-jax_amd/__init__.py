"""MI355X-native FVDBM solver with the API of MADALERAU/FVDBM-JAX's
Environment.step() path.

    import importlib
    fv = importlib.import_module("fvdbm-jax_amd")   # also importable as fvdbm_jax_amd afterwards
    mesher = fv.Mesher.from_raw(fv.meshgen.ldc_structured(32))
    mesher.enforce_ccw(); mesher.calc_mesh_properties()
    dyn = fv.D2Q9(tau=0.6, delta_t=0.01)
    cells, faces, nodes = mesher.to_env(dyn, "upwind", dim_multiplier=1)
    nodes = mesher.set_vel_node(nodes, 3, np.array([0.1, 0.0]))
    env = fv.Environment(cells, faces, nodes); env.init()
    for _ in range(1000):
        env = env.step()
"""
import sys as _sys

from . import autotune, meshgen, porous
from .containers import Cells, CustomArray, Faces, Nodes
from .dynamics import D2Q9, D2Q13, Dynamics
from .environment import Environment
from .mesher import Mesher
from .refiner import MeshRefiner

_sys.modules.setdefault("fvdbm_jax_amd", _sys.modules[__name__])

__all__ = ["Cells", "CustomArray", "D2Q9", "D2Q13", "Dynamics", "Environment", "Faces",
           "MeshRefiner", "Mesher", "Nodes", "autotune", "meshgen", "porous"]
