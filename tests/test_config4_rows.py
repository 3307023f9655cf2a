"""The 200-atom f-recipe J/K build (nbf 6000) against the oracle's exact
rows.  The GPU side is `python bench.py --config4 fp64|mixed --rows-out
FILE`, which saves the rows of the sampled shells; this test computes the
exact rows (oracle.fock_rows_exact: every quartet touching the shell, the
ill-conditioned ones in quad precision) of the one whose rows are cheapest,
the f shell of the atom farthest from the centre (about 3.7e6 quartets,
20-25 min on 6 threads), and holds each saved build to 1e-8 on them.  Runs
only when LIBINT_AMD_CONFIG4_ROWS names the saved files (comma-separated);
no GPU is needed here."""
import os

import numpy as np
import pytest

import oracle
from libint_amd import molecule

FILES = [f for f in os.environ.get("LIBINT_AMD_CONFIG4_ROWS", "").split(",") if f]
pytestmark = pytest.mark.skipif(not FILES, reason="set LIBINT_AMD_CONFIG4_ROWS to the rows saved by bench.py --config4")


def density(nbf, seed):
    rng = np.random.default_rng(seed)
    C = rng.normal(0.0, 0.1, (nbf, max(1, nbf // 5)))
    return C @ C.T


@pytest.fixture(scope="module")
def exact_rows():
    mol = molecule.make_molecule(4, natoms=200)
    shell = molecule.sampled_row_shells(mol)[-1]
    rows, G, J, K, nq, nquad = oracle.fock_rows_exact(mol.basis, density(mol.nbf, 12), [shell], 1e-12,
                                                      nthreads=int(os.environ.get("OMP_NUM_THREADS", "6")))
    print(f"oracle: shell {shell}, {nq} quartets ({nquad} in quad precision)", flush=True)
    return rows, G, J, K


@pytest.mark.timeout(3600)
def test_config4_rows_vs_exact(exact_rows):
    rows, Go, Jo, Ko = exact_rows
    saved = [np.load(f) for f in FILES]
    for f, s in zip(FILES, saved):
        at = {int(r): i for i, r in enumerate(s["rows"])}
        idx = [at[int(r)] for r in rows]
        for name, X, ref in (("G", s["G"], Go), ("J", s["J"], Jo), ("K", s["K"], Ko)):
            d = np.abs(X[idx] - ref).max()
            print(f"{f} ({s['mode'] if 'mode' in s.files else ''}): max|{name} - exact| {d:.3e} (max|{name}| {np.abs(ref).max():.3e})", flush=True)
            assert d <= 1e-8
    if len(saved) == 2:
        d = np.abs(saved[0]["Ge"] - saved[1]["Ge"]).max()
        print(f"between the two builds, every 50th row of G: max difference {d:.3e}", flush=True)
        assert d <= 1e-8
