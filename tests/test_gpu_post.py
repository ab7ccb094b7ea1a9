"""Stress recovery and Kirsch metrics on the device (postprocess.cpp:35-150).

- element stresses of the device solution equal the oracle's (oracle/cst_oracle.c
  oc_element_stresses, the reference's operation order) bit for bit, per-element modulus
  included; von Mises and sigma_zz follow element.cpp:72-81;
- Kirsch metrics on the reference's Kirsch meshes match the reference's own numbers
  (docs/verification_results.json, kept in tests/golden/golden.json): the SCF element exactly,
  the relative errors and SCF values to 1e-6 relative (the solutions agree to 1e-9 and the
  device sums in a different order).
"""
import json
import os

import numpy as np
import pytest

import cstfem_hip
from cstfem_hip import pipeline

from test_gpu_parity import GOLD, kirsch_mesh, run_device

pytestmark = pytest.mark.gpu


def golden():
    with open(os.path.join(GOLD, "golden.json")) as f:
        return json.load(f)


def kirsch_problem(lvl):
    return pipeline.Problem(mesh=kirsch_mesh(lvl), E=210e9, nu=0.3, thickness=0.01,
                            tractions=[("right", dict(kirsch=(100e6, 0.1))),
                                       ("top", dict(kirsch=(100e6, 0.1)))],
                            dirichlet=[("left_sym", 0, 0.0), ("bottom_sym", 1, 0.0)])


@pytest.mark.parametrize("lvl", [0, 1])
def test_kirsch_metrics_match_reference(dev, lvl):
    g = golden()[f"kirsch_L{lvl}"]["kirsch"]
    r = run_device(dev, kirsch_problem(lvl), rtol=1e-12)
    assert r["st"].converged
    m = dev.kirsch_metrics(100e6, 0.1)
    assert m["scf_element"] == g["scf_element"]
    for key in ("l2_rel_stress_error", "l2_rel_displacement_error", "energy_rel_error",
                "scf_x", "scf_y", "scf_numerical", "scf_analytical_centroid"):
        assert abs(m[key] - g[key]) <= 1e-6 * abs(g[key]), (key, m[key], g[key])
    assert m["scf_theoretical"] == 3.0


@pytest.mark.parametrize("case", ["kirsch_L1", "hetero_strain"])
def test_element_stresses_bitwise(dev, oracle, case):
    if case == "kirsch_L1":
        prob = kirsch_problem(1)
    else:
        prob = pipeline.plate_problem(20, 14, Lx=2.0, Ly=1.0, jitter=0.2, seed=3, E=1.0, nu=0.3,
                                      plane=cstfem_hip.PLANE_STRAIN, tx=1.0, E_sigma=1.0,
                                      E_seed=5, body=(0.0, -1.0))
    r = run_device(dev, prob, rtol=1e-10)
    m = prob.mesh
    strain = 1 if prob.plane == cstfem_hip.PLANE_STRAIN else 0
    sig, vm, szz, cxy, area = dev.element_stresses(m.num_elements)
    osig, ocxy, oarea = oracle.element_stresses(m.xy, m.tri, r["x"], prob.E, prob.nu, strain,
                                                E_elem=prob.E_elem)
    np.testing.assert_array_equal(sig, osig)
    np.testing.assert_array_equal(cxy, ocxy)
    np.testing.assert_array_equal(area, oarea)
    sz = prob.nu * (sig[:, 0] + sig[:, 1]) if strain else np.zeros(len(sig))
    np.testing.assert_array_equal(szz, sz)
    j2 = ((sig[:, 0] - sig[:, 1]) ** 2 + (sig[:, 1] - sz) ** 2 + (sz - sig[:, 0]) ** 2)
    np.testing.assert_allclose(vm, np.sqrt(0.5 * j2 + 3.0 * sig[:, 2] ** 2), rtol=1e-15)


def test_kirsch_metrics_need_uniform_modulus(dev):
    prob = pipeline.plate_problem(8, 8, E_sigma=1.0, E_seed=2)
    run_device(dev, prob)
    with pytest.raises(cstfem_hip.CstfemError):
        dev.kirsch_metrics(1.0, 0.1)
