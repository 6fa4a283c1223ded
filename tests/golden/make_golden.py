"""Writes the golden fixtures from the reference's own CPU sources.

Runs only in the development container, where oracle/_ref/libfemref.so can be built
from the reference's src/*.cpp (`make -C oracle ref`). It stores OUTPUTS only (the
reference's source never enters this repository):

  basis.npz     q_ref/q_weight/interp/grad of create_tensor_H1_lagrange for
                P = 2..10, Q = 1..12, Gauss and Gauss-Lobatto quadrature
  c1_mass.npz   the 16^3, P=3, Q=4 mass problem (jitter 0.1h seed 0, u ~ U[-1,1]
                seed 1): sha256 of the inputs, sha256 of qdata and every 97th
                qdata value, and v = M u in full
  manifest.json sha256 of every array above

Inputs (offsets, coordinates, u) come from the library's host-side mesh/RNG
functions, whose own tests pin them against independent numpy statements.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "fe-demo-cpp_amd"))

from oracle_lib import reference  # noqa: E402

import fehip  # noqa: E402  (host-only functions: no GPU needed)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def c1_inputs(n=16, P=3, Q=4):
    E = n ** 3
    nv = (n + 1) ** 3
    N = (n * (P - 1) + 1) ** 3
    return dict(n=n, P=P, Q=Q, E=E, nv=nv, N=N, offsets=fehip.box_offsets(n, n, n, P, 1),
                mesh_offsets=fehip.box_offsets(n, n, n, 2, 1), coords=fehip.box_coords(n, 0.1, 0),
                u=fehip.rng_uniform(1, N))


def main():
    ref = reference()
    if ref is None:
        sys.exit("oracle/_ref/libfemref.so is not built (make -C oracle ref)")
    manifest = {}
    basis = {}
    for mode in (0, 1):
        for P in range(2, 11):
            for Q in range(1, 13):
                if mode == 1 and Q < 2:
                    continue
                b = ref.basis(3, 1, P, Q, mode)
                for k, v in b.items():
                    basis["m%d_P%d_Q%d_%s" % (mode, P, Q, k)] = v
    np.savez_compressed(os.path.join(HERE, "basis.npz"), **basis)
    manifest["basis"] = {"count": len(basis), "sha256_all": sha(np.concatenate([basis[k] for k in sorted(basis)]))}

    c = c1_inputs()
    n, P, Q, E, nv, N = c["n"], c["P"], c["Q"], c["E"], c["nv"], c["N"]
    qdata = ref.build_qdata(3, E, 2, Q, 3 * nv, nv, c["mesh_offsets"], c["coords"])
    op = ref.mass(3, E, P, Q, 1, N, N, c["offsets"], qdata)
    v = op.apply(c["u"])
    out = dict(qdata_sample=qdata[::97].copy(), v=v)
    np.savez_compressed(os.path.join(HERE, "c1_mass.npz"), **out)
    manifest["c1"] = {"inputs": {k: sha(c[k]) for k in ("offsets", "mesh_offsets", "coords", "u")},
                      "qdata": sha(qdata), "qdata_sum": float(qdata.sum()), "v": sha(v),
                      "v_sum": float(v.sum())}
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote basis.npz, c1_mass.npz, manifest.json")


if __name__ == "__main__":
    main()
