"""Generate the committed golden vectors for the matrix-exponential path.

The reference (calaman_G) ships no data fixtures: its expm tests draw inputs
from a seeded std::mt19937 at run time and check closed forms and identities
(test/expm/expm_tests.cpp:315-336, 471-698). These fixtures add what those
cannot: exp(A) for dense random matrices from two independent sources,

  * mpmath.expm at 60 significant digits (rounded to double), n <= 24;
  * scipy.linalg.expm (Al-Mohy--Higham 2009 Pade, OpenBLAS), n in {32, 64};
  * scipy.linalg.expm_frechet (L(A, E)) for the adjoint, n in {4, 8}.

The input bits are stored next to the outputs, so the fixtures do not depend
on this generator's random stream once written. Regenerate with

    python tests/golden/gen_golden.py

Sizes are kept small so the .npz files stay a few hundred KB.
"""
from __future__ import annotations

import json
import os

import mpmath as mp
import numpy as np
import scipy
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
NORMS = (0.5, 4.0, 8.0, 100.0)


def norm1(a: np.ndarray) -> float:
    return float(np.abs(a).sum(axis=0).max())


def make_input(kind: str, n: int, target: float, seed: int) -> np.ndarray:
    """Gaussian input rescaled to the target 1-norm (the reference's
    scaled_random, test/expm/expm_tests.cpp:315-336, with a normal draw)."""
    rng = np.random.default_rng(seed)
    if kind == "real":
        a = rng.standard_normal((n, n))
    elif kind == "complex":
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    elif kind == "symmetric":
        g = rng.standard_normal((n, n))
        a = (g + g.T) / 2
    elif kind == "hermitian":
        g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        a = (g + g.conj().T) / 2
    elif kind == "antihermitian":
        g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        a = 1j * (g + g.conj().T) / 2
    else:
        raise ValueError(kind)
    nrm = norm1(a)
    if nrm > 0:
        a = a * (target / nrm)
    return np.asfortranarray(a)


def mp_expm(a: np.ndarray, dps: int) -> np.ndarray:
    with mp.workdps(dps):
        m = mp.matrix([[mp.mpc(complex(x)) if np.iscomplexobj(a) else mp.mpf(float(x))
                        for x in row] for row in a.tolist()])
        e = mp.expm(m)
        n = a.shape[0]
        if np.iscomplexobj(a):
            out = np.array([[complex(e[i, j]) for j in range(n)] for i in range(n)])
        else:
            out = np.array([[float(mp.re(e[i, j])) for j in range(n)] for i in range(n)])
    return np.asfortranarray(out)


def main() -> None:
    cases = []
    mp_arrays = {}
    sp_arrays = {}
    seed = 1000
    # mpmath truth
    for kind in ("real", "complex"):
        for n in (1, 2, 3, 5, 8, 13, 16, 24):
            for t in NORMS:
                a = make_input(kind, n, t, seed)
                e = mp_expm(a, 60)
                if n <= 8:
                    # self-consistency of the high-precision truth
                    e2 = mp_expm(a, 90)
                    assert np.array_equal(e, e2), (kind, n, t)
                key = f"c{len(cases)}"
                mp_arrays[key + "_A"] = a
                mp_arrays[key + "_E"] = e
                cases.append(dict(key=key, file="expm_mpmath.npz", kind=kind, n=n, norm1=t,
                                  seed=seed, source="mpmath.expm dps=60"))
                seed += 1
    for kind in ("symmetric", "hermitian", "antihermitian"):
        for n in (8, 16):
            for t in (4.0, 100.0):
                a = make_input(kind, n, t, seed)
                e = mp_expm(a, 60)
                key = f"c{len(cases)}"
                mp_arrays[key + "_A"] = a
                mp_arrays[key + "_E"] = e
                cases.append(dict(key=key, file="expm_mpmath.npz", kind=kind, n=n, norm1=t,
                                  seed=seed, source="mpmath.expm dps=60"))
                seed += 1
    # scipy (independent Pade variant) at larger n
    for kind, sizes in (("real", (32, 64)), ("complex", (32,))):
        for n in sizes:
            for t in (0.5, 8.0, 100.0):
                a = make_input(kind, n, t, seed)
                e = np.asfortranarray(sl.expm(a))
                key = f"c{len(cases)}"
                sp_arrays[key + "_A"] = a
                sp_arrays[key + "_E"] = e
                cases.append(dict(key=key, file="expm_scipy.npz", kind=kind, n=n, norm1=t,
                                  seed=seed, source=f"scipy.linalg.expm {scipy.__version__}"))
                seed += 1
    # Frechet derivative L(A, E) for the adjoint
    fr_arrays = {}
    fr_cases = []
    for kind in ("real", "complex"):
        for n in (4, 8):
            for t in (0.5, 8.0):
                a = make_input(kind, n, t, seed)
                ed = make_input(kind, n, 1.0, seed + 7919)
                expa, l = sl.expm_frechet(a, ed)
                key = f"f{len(fr_cases)}"
                fr_arrays[key + "_A"] = a
                fr_arrays[key + "_dir"] = ed
                fr_arrays[key + "_L"] = np.asfortranarray(l)
                fr_arrays[key + "_expA"] = np.asfortranarray(expa)
                fr_cases.append(dict(key=key, file="frechet_scipy.npz", kind=kind, n=n,
                                     norm1=t, seed=seed,
                                     source=f"scipy.linalg.expm_frechet {scipy.__version__}"))
                seed += 1
    np.savez_compressed(os.path.join(HERE, "expm_mpmath.npz"), **mp_arrays)
    np.savez_compressed(os.path.join(HERE, "expm_scipy.npz"), **sp_arrays)
    np.savez_compressed(os.path.join(HERE, "frechet_scipy.npz"), **fr_arrays)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(dict(expm=cases, frechet=fr_cases, mpmath=mp.__version__,
                       numpy=np.__version__), f, indent=1)
    print(f"{len(cases)} expm cases, {len(fr_cases)} frechet cases")


if __name__ == "__main__":
    main()
