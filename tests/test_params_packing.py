"""Term enumeration rules of extract_mmff_params (mlxmolkit/mmff_params.py:129-320), the
MMFFParams layout, the SMILES front end and the ragged device packing."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from mlxmolkit_archive_amd import molgen, packing
from mlxmolkit_archive_amd.mmff_params import MMFFParams, extract_mmff_params
from mlxmolkit_archive_amd.mol import mol_from_smiles

REF_FIELDS = ["n_atoms", "bond_idx1", "bond_idx2", "bond_kb", "bond_r0", "angle_idx1", "angle_idx2", "angle_idx3",
              "angle_ka", "angle_theta0", "angle_is_linear", "strbend_idx1", "strbend_idx2", "strbend_idx3",
              "strbend_kba_ijk", "strbend_kba_kji", "strbend_r0_ij", "strbend_r0_kj", "strbend_theta0", "oop_idx1",
              "oop_idx2", "oop_idx3", "oop_idx4", "oop_koop", "torsion_idx1", "torsion_idx2", "torsion_idx3",
              "torsion_idx4", "torsion_V1", "torsion_V2", "torsion_V3", "vdw_idx1", "vdw_idx2", "vdw_R_star",
              "vdw_eps", "ele_idx1", "ele_idx2", "ele_charge_term", "ele_diel_model", "ele_is_1_4"]


def _mol(smi):
    m = mol_from_smiles(smi)
    m.set_conformers(np.zeros((1, m.GetNumAtoms(), 3)))
    return m, extract_mmff_params(m)


def test_mmffparams_field_layout_matches_reference():
    assert [f for f in MMFFParams.__dataclass_fields__] == REF_FIELDS


@pytest.mark.parametrize("smi,n", [("CCO", 9), ("c1ccccc1", 12), ("CC(=O)Oc1ccccc1C(=O)O", 21),
                                   ("CN1C=NC2=C1C(=O)N(C(=O)N2C)C", 24), ("CCCCCCCCCCCCCCCC", 50),
                                   ("c1ccsc1", 9), ("CC#N", 6)])
def test_smiles_explicit_hydrogens(smi, n):
    assert mol_from_smiles(smi).GetNumAtoms() == n


@pytest.mark.parametrize("smi", ["CC(=O)Oc1ccccc1C(=O)O", "CCC#N", "C1CCOC1", "c1ccncc1CCl", "CC(=O)NC"])
def test_enumeration_rules(smi):
    m, p = _mol(smi)
    n = m.GetNumAtoms()
    deg = np.array([len(m.neighbors(i)) for i in range(n)])
    assert len(p.bond_idx1) == len(m.bonds)
    assert len(p.angle_idx1) == int(sum(d * (d - 1) // 2 for d in deg))
    # stretch-bend: every angle except those centred on linear (sp) atoms
    lin = {int(c) for c, l in zip(p.angle_idx2, p.angle_is_linear) if l}
    assert len(p.strbend_idx1) == sum(1 for c in p.angle_idx2 if int(c) not in lin)
    # out-of-plane: 3 permutations per trivalent centre that has parameters, centre = idx2
    for c in set(p.oop_idx2.tolist()):
        assert deg[c] == 3 and (p.oop_idx2 == c).sum() == 3
    # torsions: i1-i2-i3-i4 bonded chain, i1 != i3, i4 not in {i2, i1}
    bonds = {(min(a, b), max(a, b)) for a, b in zip(p.bond_idx1, p.bond_idx2)}
    for a, b, c, d in zip(p.torsion_idx1, p.torsion_idx2, p.torsion_idx3, p.torsion_idx4):
        assert {(min(a, b), max(a, b)), (min(b, c), max(b, c)), (min(c, d), max(c, d))} <= bonds
        assert a != c and d not in (a, b)
    # non-bonded: exactly the pairs at topological distance >= 3; 1-4 flag = distance 3
    dm = m.topological_distances()
    want = {(i, j) for i, j in itertools.combinations(range(n), 2) if dm[i, j] >= 3}
    got = set(zip(p.vdw_idx1.tolist(), p.vdw_idx2.tolist()))
    assert got == want
    for i, j, f in zip(p.ele_idx1, p.ele_idx2, p.ele_is_1_4):
        assert (dm[i, j] == 3) == bool(f) and dm[i, j] >= 3
    assert p.ele_diel_model == 0


def test_linear_centre_flag_and_nonbonded_cutoff():
    m, p = _mol("CCC#N")
    lin_centres = {int(c) for c, l in zip(p.angle_idx2, p.angle_is_linear) if l}
    assert lin_centres == {2}  # the nitrile carbon
    # 100 A cutoff on the chosen conformer drops far pairs (mmff_params.py:288-292)
    m2 = mol_from_smiles("CCCCO")
    X = np.zeros((m2.GetNumAtoms(), 3))
    X[4] = (200.0, 0, 0)
    m2.set_conformers(X[None])
    p2 = extract_mmff_params(m2)
    assert all(4 not in (i, j) for i, j in zip(p2.vdw_idx1, p2.vdw_idx2))


def test_variant_mmff94s_changes_only_conjugated_nitrogen_terms():
    m = mol_from_smiles("CC(=O)Nc1ccccc1")
    m.set_conformers(np.zeros((1, m.GetNumAtoms(), 3)))
    a = extract_mmff_params(m, mmff_variant="MMFF94")
    b = extract_mmff_params(m, mmff_variant="MMFF94s")
    assert np.array_equal(a.bond_kb, b.bond_kb) and np.array_equal(a.vdw_eps, b.vdw_eps)
    assert not np.array_equal(a.oop_koop, b.oop_koop)
    with pytest.raises(ValueError):
        extract_mmff_params(m, mmff_variant="UFF")


def test_generators_are_seeded_and_valence_complete():
    a = molgen.generate("ragged", 20, seed=5)
    b = molgen.generate("ragged", 20, seed=5)
    for x, y in zip(a, b):
        assert x.atomic_numbers().tolist() == y.atomic_numbers().tolist()
    for m in a:
        for at in m.atoms:
            if at.z == 1:
                assert len(m.neighbors(at.idx)) == 1
        m.set_conformers(np.zeros((1, m.GetNumAtoms(), 3)))
        extract_mmff_params(m)  # every generated molecule types


def test_merge_pairs_round_trip_and_order():
    m, p = _mol("CC(=O)Oc1ccccc1C(=O)O")
    i, j, R, eps, qq = packing.merge_pairs(p)
    keyv = {(a, b): (r, e) for a, b, r, e in zip(p.vdw_idx1, p.vdw_idx2, p.vdw_R_star, p.vdw_eps)}
    keye = {(a, b): q * (0.75 if f else 1.0) for a, b, q, f in zip(p.ele_idx1, p.ele_idx2, p.ele_charge_term,
                                                                   p.ele_is_1_4)}
    for a, b, r, e, q in zip(i, j, R, eps, qq):
        if (a, b) in keyv:
            assert (r, e) == keyv[(a, b)]
        assert q == pytest.approx(keye.get((a, b), 0.0), rel=1e-6)
    assert len(i) == len(set(keyv) | set(keye))
    d = j - i  # sorted by (j - i, i): consecutive lanes touch distinct atoms
    assert np.all(np.diff(d) >= 0)


def test_pack_layout_and_validation():
    ps = [_mol(s)[1] for s in ("CCO", "c1ccccc1", "CC#N")]
    hb = packing.pack(ps, [2, 1, 3])
    desc = hb.mol_desc.reshape(-1, 16)
    assert desc[:, 0].tolist() == [9, 12, 6]
    assert hb.conf_mol.tolist() == [0, 0, 1, 2, 2, 2]
    assert hb.conf_off.tolist() == [0, 27, 54, 90, 108, 126] and hb.total_coords == 144
    ij = hb.bond_ij[desc[1, 1]:desc[1, 1] + desc[1, 2]]
    assert np.array_equal(ij & 0xFFFF, ps[1].bond_idx1) and np.array_equal(ij >> 16, ps[1].bond_idx2)
    bad = _mol("CCO")[1]
    bad.torsion_idx4 = bad.torsion_idx4.copy()
    bad.torsion_idx4[0] = 99
    with pytest.raises(ValueError):
        packing.pack([bad], [1])


def test_bucket_tables_cover_every_size_and_name_known_placements():
    """The default and L-BFGS bucket tables ascend and end with an unbounded class; every
    autotuner candidate names a block the planner accepts and a known placement. The staged
    bonded-table size is the full staging size without the pair tables."""
    from mlxmolkit_archive_amd import _lib, engine
    places = {_lib.HESS_AUTO, _lib.HESS_LDS, _lib.HESS_GLOBAL, _lib.HESS_GLOBAL_PACKED, _lib.HESS_GLOBAL_SYM}
    for table in (engine.DEFAULT_BUCKETS, engine.LBFGS_BUCKETS):
        bounds = [b for b, _, _ in table]
        assert bounds == sorted(bounds) and bounds[-1] >= 3 * 65535
        for _, block, place in table:
            assert block % 64 == 0 and 64 <= block <= 1024 and place in places
    bounds = [b for b, _ in engine.TUNE_CLASSES]
    assert bounds == sorted(bounds) and bounds[-1] >= 3 * 65535
    for _, cands in engine.TUNE_CLASSES:
        for block, place in cands:
            assert block % 64 == 0 and 64 <= block <= 1024 and place in places
    params = [_mol(smi)[1] for smi in ("CCO", "c1ccccc1", "CC(=O)Nc1ccccc1", "CCCCCC")]
    hb = packing.pack(params, [1] * len(params))
    full = engine.table_lds_bytes(hb)
    bonded = engine.table_lds_bytes(hb, pairs=False)
    npair = hb.mol_desc.reshape(-1, 16)[:, 12].astype(np.int64)
    assert np.array_equal(full - bonded, 20 * npair)


def test_ignore_interfrag_interactions_opt_in():
    """Default (mmff_params.py:284-309): pairs between disconnected fragments are kept;
    ignore_interfrag_interactions=True (mmff_extract_shivam.py:361-378) drops exactly those."""
    m = mol_from_smiles("CCO.O")
    m.set_conformers(np.zeros((1, m.GetNumAtoms(), 3)))
    a = extract_mmff_params(m)
    b = extract_mmff_params(m, ignore_interfrag_interactions=True)
    frag = np.full(m.GetNumAtoms(), -1)
    dm = m.topological_distances()
    for i in range(m.GetNumAtoms()):
        if frag[i] < 0:
            frag[dm[i] < m.GetNumAtoms()] = frag.max() + 1
    assert frag.max() == 1
    cross = lambda p, q: frag[p] != frag[q]  # noqa: E731
    assert cross(a.vdw_idx1, a.vdw_idx2).any()
    assert not cross(b.vdw_idx1, b.vdw_idx2).any() and not cross(b.ele_idx1, b.ele_idx2).any()
    keep = ~cross(a.vdw_idx1, a.vdw_idx2)
    assert np.array_equal(a.vdw_idx1[keep], b.vdw_idx1) and np.array_equal(a.vdw_R_star[keep], b.vdw_R_star)
    keep_e = ~cross(a.ele_idx1, a.ele_idx2)
    assert np.array_equal(a.ele_charge_term[keep_e], b.ele_charge_term)


@pytest.mark.parametrize("smi", ["CC#N", "C=C=C", "O=C=O", "C#CC(C)C=C"])
def test_linear_rule_mmff_types_opt_in(smi):
    """linear_rule='mmff_types' flags exactly the angle centres of MMFF type 4, 53 or 61
    (mmff_extract_shivam.py:16-17,204) and drops the stretch-bends centred on them (:230-238)."""
    from mlxmolkit_archive_amd.mmff_params import MMFF_LINEAR_TYPES, _OwnSource
    m = mol_from_smiles(smi)
    m.set_conformers(np.zeros((1, m.GetNumAtoms(), 3)))
    p = extract_mmff_params(m, linear_rule="mmff_types")
    types = np.asarray(_OwnSource(m, "MMFF94").t.types)
    want = np.isin(types[p.angle_idx2], list(MMFF_LINEAR_TYPES)).astype(np.int32)
    assert np.array_equal(p.angle_is_linear, want)
    assert not np.isin(types[p.strbend_idx2], list(MMFF_LINEAR_TYPES)).any()
    with pytest.raises(ValueError):
        extract_mmff_params(m, linear_rule="bogus")
