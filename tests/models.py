"""Model descriptions shared by the oracle and GPU tests.

examples_fast: the reference's own test model,
xpsi/tests/test_likelihood.py:21-134 and
examples/examples_fast/Modules/main.py:15-155.
"""
import math

import numpy as np


def examples_fast_support(bg_counts, exposure=1000.0):
    """Background support recipe of examples/examples_fast/Modules/main.py:65-82."""
    bg = bg_counts / exposure
    sup = np.ones((len(bg), 2))
    sup[:, 0] = 0.9 * bg
    sup[:, 1] = 1.1 * bg
    for i in range(sup.shape[0]):
        if sup[i, 1] == 0.0:
            for j in range(1, sup.shape[0]):
                if i + j < sup.shape[0] and sup[i + j, 1] > 0.0:
                    sup[i, 1] = sup[i + j, 1]
                    break
                elif i - j >= 0 and sup[i - j, 1] > 0.0:
                    sup[i, 1] = sup[i - j, 1]
                    break
    return sup


def examples_fast_model(fx, radius=10.0, logT=6.8, num_energies=128,
                        support=None, mass=1.4, cos_i=None, colat=None,
                        spot_radius=0.75, distance=1.0, phase_shift=0.0):
    return dict(
        spacetime=dict(mass=mass, radius=radius, distance=distance,
                       cos_inclination=math.cos(60 * np.pi / 180) if cos_i is None else cos_i,
                       frequency=314.0, star_shape="AGM_14"),
        hot_regions=[dict(super_colatitude=70 * np.pi / 180 if colat is None else colat,
                          super_radius=spot_radius, super_temperature=logT,
                          phase_shift=phase_shift, sqrt_num_cells=32,
                          min_sqrt_num_cells=16, max_sqrt_num_cells=64,
                          num_leaves=64, num_rays=512, is_antiphased=True,
                          image_order_limit=3)],
        instrument=dict(matrix=np.diag(fx["matrix_diag"]),
                        energy_edges=fx["energy_edges"]),
        data=dict(counts=fx["counts"], phases=np.linspace(0, 1, 33),
                  exposure_time=1000.0),
        num_energies=num_energies, support=support)
