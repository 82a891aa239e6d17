"""
Distance priors of the analysis (src/exmarg/priors.py:24-49).

  'volume'          p(y) = 3 ymin^3 / y^4 for y >= ymin, integrated in closed form
  'madau-dickinson' the density of ln y of mergers whose rate follows the Madau-Dickinson history to
                    z = 10: the tabulated p(ln y) of the reference population
                    (src/exmarg/data/demo_population.npz, members lny, p_lny, ymin), kept here as data

population_m1_z() gives the population itself (src/exmarg/priors.py:41-44): the source-frame primary
mass density (GWTC-5 broken power law with two peaks) and the Madau-Dickinson redshift density on
their grids, the members m1, p_m1, z, p_z of the same reference data file (data/population.npz).
"""

import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "distance_prior.npz")
_cache = {}


def _population():
    if "pop" not in _cache:
        with np.load(_DATA) as z:
            _cache["pop"] = (np.array(z["lny"]), np.array(z["p_lny"]), float(z["ymin"]))
    return _cache["pop"]


def distance_density(name):
    """None for 'volume' (closed form), (lny, p_lny) for 'madau-dickinson' (src/exmarg/priors.py:31-38)"""
    if name == "volume":
        return None
    if name == "madau-dickinson":
        lny, p, _ = _population()
        return lny, p
    raise ValueError("the distance prior is 'volume' or 'madau-dickinson', not %r" % (name,))


def population_m1_z():
    """((m1, p_m1), (z, p_z)): source-frame primary mass and redshift densities of the population, each
    normalized on its grid (src/exmarg/priors.py:41-44)"""
    if "m1z" not in _cache:
        with np.load(os.path.join(os.path.dirname(_DATA), "population.npz")) as z:
            _cache["m1z"] = ((np.array(z["m1"]), np.array(z["p_m1"])), (np.array(z["z"]), np.array(z["p_z"])))
    return _cache["m1z"]


def madau_dickinson_ymin():
    """the smallest y = D0 / D of the Madau-Dickinson prior, the distance at z = 10 (src/exmarg/priors.py:47-49)"""
    return _population()[2]
