"""PDB input: ATOM / HETATM records -> bodies with united-atom form-factor types
(resaxslib/src/parse_pdb.cpp:35-364, typing tables resaxslib/src/utils.cpp:201-314).

The expected types below are entries of the reference's typing tables (element map,
label defaults, residue exceptions), checked one by one; record handling follows
parse_pdb.cpp (waters, hydrogens, alternate locations, models).  Host code: no GPU.
"""
import numpy as np
import pytest

# united-atom type indices (the factor table's columns)
H, C, N, O, P, S, Zn, CH, CH2, CH3, NH, NH2, NH3, OH, OH2, SH = range(16)


def atom(rec, serial, name, res, x, y, z, element="", alt=" ", chain="A", resseq=1):
    """One fixed-column PDB record (names of 1-letter elements start in column 14)."""
    nm = f" {name:<3}" if len(name) < 4 else name
    line = (f"{rec:<6}{serial:>5} {nm:<4}{alt:1}{res:>3} {chain:1}{resseq:>4}    "
            f"{x:8.3f}{y:8.3f}{z:8.3f}{1.0:6.2f}{0.0:6.2f}")
    return line + (f"          {element:>2}  " if element else "")


def test_typing_table_entries():
    from resaxs_amd import pdb_atom_type as t
    cases = [("CA", "GLY", "C", CH2), ("CA", "ALA", "C", CH), ("CB", "ALA", "C", CH3), ("CB", "SER", "C", CH2),
             ("CB", "VAL", "C", CH), ("CG", "GLU", "C", CH2), ("CG", "ASP", "C", C), ("CG", "LEU", "C", CH),
             ("CG1", "VAL", "C", CH3), ("CG1", "ILE", "C", CH2), ("CG2", "ILE", "C", CH3), ("CD", "PRO", "C", CH2),
             ("CD", "GLN", "C", C), ("CE", "MET", "C", CH3), ("CZ", "PHE", "C", CH), ("C", "ALA", "C", C),
             ("N", "ALA", "N", NH), ("N", "PRO", "N", N), ("NZ", "LYS", "N", NH3), ("ND2", "ASN", "N", NH2),
             ("NE", "ARG", "N", NH), ("O", "ALA", "O", O), ("OG", "SER", "O", OH), ("OH", "TYR", "O", OH),
             ("OD1", "ASP", "O", O), ("SG", "CYS", "S", SH), ("SD", "MET", "S", S), ("P", "DA", "P", P),
             ("ZN", "ZN", "Zn", Zn)]
    for label, res, el, want in cases:
        assert t(label, res, el) == want, (label, res, el)
    assert t("CA", "GLY") == CH2                      # element from the label's first letter
    assert t("FE", "HEM", "FE") == N                  # unknown element: N
    assert t("CB", "ALA", C) == CH3 and t("OG", "SER", O) == OH   # the element-type overload
    assert t("XX", "UNK", CH2) == CH2


def test_read_records_models_and_skips(tmp_path):
    from resaxs_amd import RsxError, read_pdb
    lines = [
        "HEADER    SYNTHETIC",
        "MODEL        1",
        atom("ATOM", 1, "N", "ALA", 1.0, 2.0, 3.0, "N"),
        atom("ATOM", 2, "CA", "ALA", 1.5, 2.5, 3.5, "C"),
        atom("ATOM", 3, "CB", "ALA", -1.25, 0.125, 10.5, "C"),
        atom("ATOM", 4, "H", "ALA", 9.0, 9.0, 9.0, "H"),           # hydrogen: skipped
        atom("ATOM", 5, "CA", "GLY", 4.0, 5.0, 6.0, "C", alt="A"),  # alt A kept
        atom("ATOM", 6, "CA", "GLY", 4.1, 5.1, 6.1, "C", alt="B"),  # alt B skipped
        atom("HETATM", 7, "O", "HOH", 0.0, 0.0, 0.0, "O"),          # water skipped
        atom("HETATM", 8, "FE", "HEM", 7.0, -8.0, 9.25, "FE"),       # unknown element: N
        atom("ATOM", 9, "SG", "CYS", -100.5, 200.25, -300.125),     # no element column: label
        "ENDMDL",
        "MODEL        2",
        atom("ATOM", 1, "OG", "SER", 11.0, 12.0, 13.0, "O"),
        atom("ATOM", 2, "NZ", "LYS", 14.0, 15.0, 16.0, "N"),
        "ENDMDL",
        "END",
    ]
    path = tmp_path / "two_models.pdb"
    path.write_text("\n".join(lines) + "\n")
    b, off = read_pdb(str(path), all_models=True)
    assert list(off) == [0, 6, 8]
    types = b[:, 3].view(np.uint32)
    assert list(types) == [NH, CH, CH3, CH2, N, SH, OH, NH3]
    expect = np.array([[1, 2, 3], [1.5, 2.5, 3.5], [-1.25, 0.125, 10.5], [4, 5, 6], [7, -8, 9.25],
                       [-100.5, 200.25, -300.125], [11, 12, 13], [14, 15, 16]], dtype=np.float32)
    assert np.array_equal(b[:, :3], expect)
    b1, off1 = read_pdb(str(path), all_models=False)   # the first model only
    assert list(off1) == [0, 6] and np.array_equal(b1, b[:6])
    with pytest.raises(RsxError) as ex:
        read_pdb(str(tmp_path / "missing.pdb"))
    assert ex.value.code == -1100


def test_records_outside_models(tmp_path):
    """Atoms before any MODEL record start a model; END stops reading; a last line
    without a newline is dropped by the all-models read (parse_pdb.cpp:72-73)."""
    from resaxs_amd import read_pdb
    lines = [atom("ATOM", 1, "C", "GLY", 1.0, 1.0, 1.0, "C"), atom("ATOM", 2, "O", "GLY", 2.0, 2.0, 2.0, "O"),
             "END", atom("ATOM", 3, "N", "GLY", 3.0, 3.0, 3.0, "N")]
    path = tmp_path / "plain.pdb"
    path.write_text("\n".join(lines) + "\n")
    b, off = read_pdb(str(path))
    assert list(off) == [0, 2] and list(b[:, 3].view(np.uint32)) == [C, O]
    path.write_text("\n".join(lines[:2]))          # no newline after the last record
    b, off = read_pdb(str(path))
    assert list(off) == [0, 1]
    b, off = read_pdb(str(path), all_models=False)
    assert list(off) == [0, 2]


def test_calc_saxs_driver_reads_pdb_models(tmp_path):
    """The C++ fitting driver built from a PDB file the tool's way: each MODEL is one
    ensemble member (calc_saxs.cpp:279-303); a missing file is resaxs error -1100."""
    import os
    import subprocess
    binary = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp", "fit_test")
    assert os.path.exists(binary), "build tests/cpp/fit_test first (__graft_entry__.build())"
    lines = ["MODEL        1"] + [atom("ATOM", i + 1, "CA", "ALA", i, 0.0, 0.0, "C") for i in range(5)] + ["ENDMDL"]
    lines += ["MODEL        2"] + [atom("ATOM", i + 1, "CA", "ALA", 0.0, i, 0.0, "C") for i in range(5)]
    lines += ["ENDMDL", "END"]
    path = tmp_path / "ens.pdb"
    path.write_text("\n".join(lines) + "\n")
    r = subprocess.run([binary, "pdb", str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "models 2 atoms 5 q 50" in r.stdout
    r = subprocess.run([binary, "pdb", str(tmp_path / "none.pdb")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "error -1100" in r.stdout
