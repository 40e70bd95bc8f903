"""Writes the regression fixtures of tests/test_golden.py (and, under
variants/, those of tests/test_golden_variants.py).

Each fixture holds a small seeded scene (inputs) and the CPU oracle's
outputs for the five calls of the contact step.  Regenerate with
`python tests/golden/make_golden.py` after an intended change of the
oracle; the KAT tests pin the oracle itself against the reference's own
test values."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "ipc-toolkit_amd"), os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from ipctk_amd import scenes  # noqa: E402


def canon(a):
    a = np.asarray(a, dtype=np.int32).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))] if len(a) else a


def fixture(name, s, per_candidate_toi):
    ee, fv = O.candidates(s.V, s.E, s.F, 0.5 * s.dhat)
    c = O.build_collisions(s.V, s.E, s.F, ee, fv, s.dhat)
    H = O.barrier_hessian(c, s.V, s.E, s.F, s.dhat, s.kappa, "clamp")
    see, sfv = O.candidates(s.V, s.E, s.F, 0.0, V1=s.V1)
    see, sfv = canon(see), canon(sfv)
    out = dict(
        V=s.V, V1=s.V1, E=s.E, F=s.F, dhat=s.dhat, kappa=s.kappa,
        cand_ee=canon(ee), cand_fv=canon(fv),
        vv=c.vv, vv_w=c.vv_w, ev=c.ev, ev_w=c.ev_w, ee=c.ee, ee_w=c.ee_w,
        ee_eps_x=c.ee_eps_x, ee_dtype=c.ee_dtype, fv=c.fv, fv_w=c.fv_w,
        energy=O.barrier_energy(c, s.V, s.E, s.F, s.dhat, s.kappa),
        gradient=O.barrier_gradient(c, s.V, s.E, s.F, s.dhat, s.kappa),
        H_indptr=H.indptr.astype(np.int64), H_indices=H.indices.astype(np.int32), H_data=H.data,
        swept_ee=see, swept_fv=sfv,
        stepsize=O.compute_collision_free_stepsize(s.V, s.V1, s.E, s.F),
    )
    if per_candidate_toi:
        out["toi_ee"] = O.candidates_toi(s.V, s.V1, s.E, s.F, see, True)
        out["toi_fv"] = O.candidates_toi(s.V, s.V1, s.E, s.F, sfv, False)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, s.sizes, "collisions", len(c), "nnz", H.nnz, "swept", len(see) + len(sfv))


def variants(name, s, mult=3.0):
    """collision-set variants and friction on one scene (tests/golden/variants/)"""
    dhat = mult * s.dhat
    ee, fv = O.candidates(s.V, s.E, s.F, 0.5 * dhat)
    out = dict(V=s.V, E=s.E, F=s.F, dhat=dhat, kappa=s.kappa)
    for tag, kw in (("imp", dict(area_weighting=True, improved=True)), ("ogc", dict(ogc=True))):
        c = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat, **kw)
        for k in ("vv", "ev", "ee", "fv", "vv_w", "ev_w", "ee_w", "fv_w", "ee_eps_x", "ee_dtype"):
            out[f"{tag}_{k}"] = getattr(c, k)
    c = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat)
    fr = O.friction_build(c, s.V, s.E, s.F, dhat, s.kappa, 0.6, 0.3)
    rng = np.random.default_rng(11)
    n = len(fr.keep)
    aniso = dict(mu_aniso=rng.uniform(0.5, 2.0, (n, 2)), mu_s_aniso=rng.uniform(0.2, 0.8, (n, 2)),
                 mu_k_aniso=rng.uniform(0.1, 0.5, (n, 2)))
    Vel = rng.normal(0, 2e-3, s.V.shape)
    O.friction_set_anisotropy(fr, **aniso)
    O.friction_update_lagged_mu(fr, Vel)
    H = O.friction_hessian(fr, Vel, 1e-3, "clamp")
    out.update(Vel=Vel, eps_v=1e-3, **aniso,
               fric_energy=O.friction_energy(fr, Vel, 1e-3),
               fric_gradient=O.friction_gradient(fr, Vel, 1e-3),
               fric_H_indptr=H.indptr.astype(np.int64), fric_H_indices=H.indices.astype(np.int32),
               fric_H_data=H.data)
    os.makedirs(os.path.join(HERE, "variants"), exist_ok=True)
    np.savez_compressed(os.path.join(HERE, "variants", name + ".npz"), **out)
    print("variants", name, "improved", len(out["imp_ee"]), "ogc", len(out["ogc_fv"]), "nnz", H.nnz)


if __name__ == "__main__":
    O.build()
    variants("twisted_sheets_8_seed0", scenes.two_sheets(8, seed=0, twist=0.3))
    fixture("twisted_sheets_8_seed0", scenes.two_sheets(8, seed=0, twist=0.3), False)
    fixture("sphere_chain_4x3_seed0", scenes.sphere_chain(4, 3, seed=0), True)
    fixture("stacked_sheets_4x12_seed2", scenes.stacked_sheets(12, 4, seed=2, h=1.0 / 511), True)
