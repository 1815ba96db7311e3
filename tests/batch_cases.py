"""Sampled parameter vectors of the batched workload for parity tests."""
import math

from xpsi_amd import configs


def oracle_model(d, ne):
    """The oracle's model dict for one config-3 draw."""
    m = configs.config3_model(ne)
    m["spacetime"].update(mass=d["mass"], radius=d["radius"], distance=d["distance"],
                          cos_inclination=d["cos_inclination"])
    h = m["hot_regions"][0]
    hp = d["hot_regions"][0]
    h.update(super_colatitude=hp["super_colatitude"], super_radius=hp["super_radius"],
             super_temperature=hp["super_temperature"], phase_shift=hp["phase_shift"],
             sqrt_num_cells=hp["sqrt_num_cells"])
    return m


def sample(n_take=12, max_cells=48, max_e=128):
    draws = configs.config3_draws()
    out = []
    for d in draws:
        if not configs.prior_accepts(d["mass"], d["radius"], 300.0):
            continue
        if d["hot_regions"][0]["sqrt_num_cells"] > max_cells or d["num_energies"] > max_e:
            continue
        out.append(d)
        if len(out) >= n_take:
            break
    return out


def is_polar(d):
    h = d["hot_regions"][0]
    return h["super_colatitude"] - h["super_radius"] < 0 or h["super_colatitude"] + h["super_radius"] > math.pi
