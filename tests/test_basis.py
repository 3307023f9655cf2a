"""libint_amd.basis: Gaussian94 reading and BasisSet assembly with
libint2::BasisSet's rules (include/libint2/basis.h.in), .xyz reading with
read_dotxyz's (include/libint2/atom.h).  Synthetic files in both g94
layouts; and, where the reference's data files are present in this
container, aug-cc-pVDZ water read through BasisSet equals the committed
fixture that tests/golden/make_h2o_augccpvdz.py wrote with its own parser."""
import json
import os

import numpy as np
import pytest

import libint_amd as la
from libint_amd import basis

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = "/root/reference/lib"

EMSL = """! a comment
****
H     0
S   2   1.00
      3.42525091D+00         0.15432897D+00
      0.62391373D+00         0.53532814D+00
****
C     0
SP   2   1.00
      2.9412494             -0.09996723             0.15591627
      0.6834831              0.39951283             0.60768372
D   1   1.00
      0.8                    1.0
F   1   1.00
      0.7                    1.0
****
"""

MOLSSI = """H     0
S   1   1.00
      0.1220000              1.0000000
P   1   1.00
      0.7270000              1.0000000
****
"""


def test_g94_layouts_sp_and_purity(tmp_path):
    (tmp_path / "mybasis.g94").write_text(EMSL)
    lib = basis.read_g94(str(tmp_path / "mybasis.g94"))
    assert [s[0] for s in lib[1]] == [0]
    assert lib[1][0][2] == [3.42525091, 0.62391373]
    # SP split into s then p (Cartesian), d and f pure by default
    assert [(l, p) for l, p, _, _ in lib[6]] == [(0, False), (1, False), (2, True), (3, True)]
    assert lib[6][1][3] == [0.15591627, 0.60768372]
    cart = basis.read_g94(str(tmp_path / "mybasis.g94"), force_cartesian_d=True)
    assert [(l, p) for l, p, _, _ in cart[6]] == [(0, False), (1, False), (2, False), (3, True)]
    (tmp_path / "new.g94").write_text(MOLSSI)
    assert [s[0] for s in basis.read_g94(str(tmp_path / "new.g94"))[1]] == [0, 1]
    with pytest.raises(OSError):
        basis.read_g94(str(tmp_path / "missing.g94"))
    assert basis.read_g94(str(tmp_path / "missing.g94"), throw_if_missing=False) == {}


def test_names_and_conventions():
    assert basis.canonicalize_name("6-31G*") == "6-31g*"
    assert basis.canonicalize_name("def2-SVP/JK") == "def2-svpIjk"
    for name, cart in (("6-31g*", True), ("6-31g", True), ("6-31+g*", True), ("6-311g**", False),
                       ("3-21g", True), ("4-31g", True), ("cc-pvdz", False), ("6-31g(d)", False)):
        assert basis.gaussian_cartesian_d_convention(name) == cart, name
    assert basis.decompose_name_into_components("aug-cc-pvdz") == ["cc-pvdz", "augmentation-cc-pvdz"]
    assert basis.decompose_name_into_components("aug-cc-pvdz-cabs") == ["aug-cc-pvdz-cabs"]


def test_basisset_assembly(tmp_path, monkeypatch):
    d = tmp_path / "basis"
    d.mkdir()
    (d / "cc-pvdz.g94").write_text(EMSL)
    (d / "augmentation-cc-pvdz.g94").write_text(MOLSSI)
    monkeypatch.setenv("LIBINT_DATA_PATH", str(tmp_path))
    atoms = basis.read_dotxyz("2\nCH\nC 0 0 0\nH 0 0 1.0\n")
    assert atoms[1].atomic_number == 1 and abs(atoms[1].z - 1.0 / 0.529177210903) < 1e-14
    bs = basis.BasisSet("aug-cc-pVDZ", atoms)
    # C from cc-pvdz only (no C in the augmentation file), then H: cc-pvdz s, augmentation s, p
    assert [s.l for s in bs] == [0, 1, 2, 3, 0, 0, 1]
    assert bs.shell2atom() == [0, 0, 0, 0, 1, 1, 1]
    assert bs.nbf() == 1 + 3 + 5 + 7 + 1 + 1 + 3 and bs.shell2bf()[4] == 16
    assert bs.max_l() == 3 and bs.max_nprim() == 2
    ref = la.Shell(0, [(3.42525091, 0.15432897), (0.62391373, 0.53532814)], (0, 0, atoms[1].z))
    assert np.allclose(bs[4].coeff, ref.coeff, rtol=0, atol=1e-15) and np.array_equal(bs[4].O, ref.O)
    with pytest.raises(LookupError):
        basis.BasisSet("aug-cc-pVDZ", atoms, throw_if_no_match=True)


@pytest.mark.skipif(not os.path.isdir(REF_LIB), reason="the reference's basis library is not present")
def test_aug_cc_pvdz_water_matches_fixture():
    d = json.load(open(os.path.join(HERE, "golden", "h2o_augccpvdz.json")))
    xyz = "3\n\n" + "\n".join(f"{el} {x} {y} {z}" for el, (x, y, z) in d["atoms_angstrom"]) + "\n"
    atoms = basis.read_dotxyz(xyz)
    bs = basis.BasisSet("aug-cc-pVDZ", atoms, throw_if_no_match=True, data_path_override=REF_LIB)
    expect = []
    for el, pos in d["atoms_angstrom"]:
        for s in d["basis"][el]:
            expect.append(la.Shell(s["l"], list(zip(s["exponents"], s["coefficients"])),
                                   tuple(np.array(pos) / 0.529177210903), pure=s["l"] > 1))
    assert len(bs) == len(expect) == 19 and bs.nbf() == 41
    for a, b in zip(bs, expect):
        assert a.l == b.l and a.pure == b.pure
        assert np.array_equal(a.alpha, b.alpha) and np.array_equal(a.coeff, b.coeff) and np.array_equal(a.O, b.O)
