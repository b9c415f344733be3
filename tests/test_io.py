"""VTK export (legacy ASCII) on the CPU with a stand-in environment."""
import types

import numpy as np

from cases import build_mesher, meshgen


def test_vtk_export(tmp_path):
    m = build_mesher(meshgen.ldc_structured(4))
    C = len(m.cells)
    cells = types.SimpleNamespace(vel=np.ones((C, 2), np.float32) * [0.1, -0.2],
                                  rho=np.full((C, 1), 1.01, np.float32),
                                  pdf=np.zeros((C, 9), np.float32), pdf_eq=np.zeros((C, 9), np.float32))
    env = types.SimpleNamespace(cells=cells)
    path = m.to_vtk(env, str(tmp_path / "ldc"), save_f=True)
    text = open(path).read().split("\n")
    assert text[0].startswith("# vtk DataFile")
    assert f"POINTS {len(m.points)} double" in text
    assert f"CELLS {C} {4 * C}" in text
    i = text.index("VECTORS Velocity double")
    assert np.allclose([float(v) for v in text[i + 1].split()], [0.1, -0.2, 0.0], atol=1e-7)
    assert "SCALARS Density double 1" in text
    assert any(line.startswith("pdf 9 ") for line in text)
