#!/usr/bin/env python3
"""Writes the irregular-mesh fixture: a permuted L-shaped 4^3-box mesh (48 hexes,
P=3) as tests/golden/lshape_n4_p3.femh, and beside it lshape_n4_p3.npz with
its offsets, the oracle's qdata (Q=4), u ~ U[-1,1] (splitmix64 seed 1) and the
oracle's v = M u. Run from the repository root after building the library and
the oracle."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "fe-demo-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import fehip  # noqa: E402
from fehip.meshes import l_shape, permute  # noqa: E402
from oracle_lib import oracle  # noqa: E402

m = permute(l_shape(4, 3), seed=11)
fehip.mesh_write(os.path.join(HERE, "lshape_n4_p3.femh"), m)
o = oracle()
E, N, nv = m["num_elem"], m["num_nodes"], m["num_vertices"]
qd = o.build_qdata(3, E, 2, 4, 3 * nv, nv, m["vertex_offsets"], m["coords"])
u = fehip.rng_uniform(1, N)
v = o.mass(3, E, 3, 4, 1, N, N, m["offsets"], qd).apply(u)
np.savez(os.path.join(HERE, "lshape_n4_p3.npz"), offsets=m["offsets"], qdata=qd, u=u, v=v)
print("wrote", E, "elements,", N, "nodes")
