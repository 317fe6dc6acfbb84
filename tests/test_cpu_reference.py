"""The oracle and the library's host-side generators against the reference's
own CPU code (oracle/_ref/libresaxs_ref.so, compiled from the reference's
sources by `make -C oracle ref`).  Skipped where that library was not built.

What this pins, bit for bit:
  host_debye<double, float, cl_float4>::calc_curve (host_debye.cpp:54-109)
      == the oracle's restatement orc_debye_ref_rows, rounded to FP32;
  atomic_form_factors<float>::generate / calc_radii / calc_avg_radius
      (ff_coef.hpp:112-196) == orc_atomic_ff / orc_atomic_radii and the
      library's rsx_atomic_form_factors / rsx_atomic_radii, expansion c != 1 included;
  create_fibonacci_dots / create_dodeca_icosa_dots (rsasa_algorithm.cpp:96-208)
      == orc_*_dots and the library's generators;
  rsasa_cpu_alg::calc_rsasa (rsasa_cpu_alg.cpp:112-185, voxel-pruned) exposed
      counts == the oracle's brute-force counts over the same dot spheres;
  parse_pdb<float> (parse_pdb.cpp:35-104, 302-364) and map_pdb_atom_to_ff_type
      (utils.cpp:201-314) == the library's rsx_pdb_read / rsx_pdb_atom_type.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "test1888.npz")

from oracle import pyoracle as po  # noqa: E402
from helpers import padded_density_dots  # noqa: E402

pytestmark = pytest.mark.skipif(po.ref_lib() is None, reason="oracle/_ref/libresaxs_ref.so not built "
                                                               "(needs the reference tree)")


def test_host_debye_equals_restatement():
    g = np.load(GOLDEN)
    ref = po.ref_host_debye(g["q"], g["pos"], g["types"], g["table"])
    mine = po.debye_ref_rows(g["q"], g["pos"], g["types"], g["table"], nthreads=4)
    assert np.array_equal(ref, mine.astype(np.float32))
    # and against the FP64 oracle (the reference rounds its FP64 sum to FP32)
    F = po.body_factors(g["pos"], g["types"], g["q"], g["table"])
    I = po.debye(g["q"], g["pos"], F)
    assert np.max(np.abs(ref - I) / I) < 1.2e-7
    # the fixture (positions rounded to 1e-3 A) to its stated accuracy
    assert np.max(np.abs(ref - g["Iq"]) / g["Iq"]) < 5e-5


def test_host_debye_random_and_edge_sizes():
    rng = np.random.default_rng(31)
    q = np.linspace(0.0, 0.5, 33).astype(np.float32)
    table = rng.uniform(1, 9, size=(33, 4)).astype(np.float32)
    for n in (1, 2, 3, 64, 65, 300):
        pos = rng.uniform(-9, 9, size=(n, 3)).astype(np.float32)
        types = rng.integers(0, 4, size=n).astype(np.uint32)
        ref = po.ref_host_debye(q, pos, types, table)
        mine = po.debye_ref_rows(q, pos, types, table, nthreads=2).astype(np.float32)
        assert np.array_equal(ref, mine), n


def test_form_factors_and_radii_bit_exact():
    import resaxs_amd as ra
    q = np.linspace(0.0, 0.75, 101).astype(np.float32)
    for r, c in ((1.62, 1.0), (1.62, 0.95), (1.58, 1.08), (1.71, 1.02)):
        ref = po.ref_atomic_ff(q, r, c)
        assert np.array_equal(ref, po.atomic_ff(q, r, c)), (r, c)
        assert np.array_equal(ref, ra.atomic_form_factors(q, r, c)), (r, c)
    radii = po.ref_atomic_radii()
    assert np.array_equal(radii, po.atomic_radii())
    assert np.array_equal(radii, ra.atomic_radii())
    # average radius over bodies of mixed types (ff_coef.hpp:186-194)
    rng = np.random.default_rng(4)
    types = rng.integers(0, 16, size=777).astype(np.uint32)
    b = po.bodies(rng.normal(size=(777, 3)), types)
    acc = np.float32(0)
    for t in types:
        acc = np.float32(acc + radii[t])
    assert po.ref_avg_radius(b) == np.float32(acc / np.float32(777))


def test_dot_generators_bit_exact():
    import resaxs_amd as ra
    for r in (1.07, 1.2, 1.58, 1.7, 1.97):
        for p in (0.0, 1.4, 1.8):
            fib = po.ref_fibonacci_dots(r, p)
            assert fib.shape == (64, 4)
            assert np.array_equal(fib, po.fibonacci_dots(32, r, p))
            assert np.array_equal(fib, ra.fibonacci_dots(32, r, p))
            ico = po.ref_dodeca_icosa_dots(r, p)
            assert ico.shape == (32, 4)
            assert np.array_equal(ico, po.dodeca_icosa_dots(r, p))
            assert np.array_equal(ico, ra.dodeca_icosa_dots(r, p))


def test_sphere_dot_generator_bit_exact():
    """The library's rings-of-dots generator equals create_dots / create_32_dots."""
    import resaxs_amd as ra
    for r in (1.07, 1.2, 1.58, 1.7, 1.97, 2.3):
        for p in (0.0, 1.4, 1.8):
            for dens in (0.25, 0.5, 2.0, 5.0):
                assert np.array_equal(ra.density_dots(dens, r, p), po.ref_density_dots(dens, r, p)), (r, p, dens)
    assert len(ra.sphere_dots(1.7, 10.5, 1.4)) > 0


def test_rsasa_cpu_counts_equal_oracle():
    radii = po.ref_atomic_radii()
    probe = 1.8
    rng = np.random.default_rng(3)
    for density in (0.5, 2.0):
        dots, nd = padded_density_dots(density, radii, probe)
        R2 = ((radii + np.float32(probe)) ** 2).astype(np.float32)
        for n, half in ((1, 1.0), (60, 8.0), (400, 14.0), (1500, 22.0)):
            pos = rng.uniform(-half, half, size=(n, 3)).astype(np.float32)
            types = rng.integers(0, 16, size=n).astype(np.uint32)
            frac = po.ref_rsasa_cpu(po.bodies(pos, types), density, probe)
            ref_counts = np.rint(frac * nd[types]).astype(np.int64)
            assert np.allclose(ref_counts / nd[types], frac, rtol=0, atol=1e-12)
            got = po.sasa_counts(pos, types, R2, dots, cells=False).astype(np.int64)
            assert np.array_equal(ref_counts, got), (density, n)


# residues and atom labels whose cross product covers the typing tables' default labels and
# every residue exception, plus labels and residues the tables do not name
PDB_RESIDUES = ["ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE", "LEU", "LYS", "MET", "PHE",
                "PRO", "SER", "THR", "TRP", "TYR", "VAL", "DA", "HEM", "UNK", "ZN"]
PDB_LABELS = ["C", "CA", "CB", "CG", "CG1", "CG2", "CD", "CD1", "CD2", "CE", "CE1", "CE2", "CE3", "CZ", "CZ2", "CZ3",
              "CH2", "CH3", "N", "ND1", "ND2", "NE", "NE1", "NE2", "NH1", "NH2", "NZ", "O", "OD1", "OD2", "OE1", "OE2",
              "OG", "OG1", "OH", "OXT", "SD", "SG", "P", "ZN", "FE", "H", "HA", "SE"]


def _pdb_line(rec, serial, name, res, x, y, z, element=None, alt=" ", chain="A", resseq=1):
    nm = f" {name:<3}" if len(name) < 4 else name
    line = (f"{rec:<6}{serial:>5} {nm:<4}{alt:1}{res:>3} {chain:1}{resseq:>4}    "
            f"{x:8.3f}{y:8.3f}{z:8.3f}{1.0:6.2f}{0.0:6.2f}")
    return line if element is None else line + f"          {element:>2}  "


def test_pdb_atom_typing_equals_reference():
    from resaxs_amd import pdb_atom_type
    for res in PDB_RESIDUES:
        for label in PDB_LABELS:
            assert pdb_atom_type(label, res) == po.ref_pdb_atom_type(label, res), (label, res)
            for el in ("H", "C", "N", "O", "P", "S", "Zn", "ZN", "FE", "Se", label[:1]):
                assert pdb_atom_type(label, res, el) == po.ref_pdb_atom_type(label, res, el), (label, res, el)


def test_pdb_reader_equals_reference(tmp_path):
    """Synthetic PDB text: two MODELs, records before the first MODEL, every residue x label of
    the typing cross product (with and without an element column, ATOM and HETATM), alternate
    locations blank / A / B / C, waters, hydrogens, unknown and two-letter elements, an END
    with records after it, a line over 99 characters and a last line without a newline.
    Records are padded to the format's 80 columns: the reference takes a record's name from
    6 bytes of its line buffer (parse_pdb.cpp:47, 88), which for a line shorter than that
    includes bytes of the previous line, so an unpadded END after ENDMDL is no END there."""
    from resaxs_amd import read_pdb
    rng = np.random.default_rng(404)
    lines, serial = [], 0

    def rec(*args, **kw):
        nonlocal serial
        serial += 1
        x, y, z = rng.uniform(-99, 99, 3)
        lines.append(_pdb_line(args[0], serial, args[1], args[2], x, y, z, **kw))

    rec("ATOM", "CA", "GLY", element="C")                     # before any MODEL
    lines.append("MODEL        1")
    for res in PDB_RESIDUES:
        for k, label in enumerate(PDB_LABELS):
            el = None if k % 3 == 0 else ("ZN" if label == "ZN" else label[:1])
            rec("HETATM" if k % 7 == 0 else "ATOM", label, res, element=el,
                alt=" ABC"[k % 4] if k % 5 == 0 else " ")
    rec("HETATM", "O", "HOH", element="O")
    rec("ATOM", "FE", "HEM", element="FE")
    lines.append("ENDMDL")
    lines.append("MODEL        2")
    for label in PDB_LABELS[:12]:
        rec("ATOM", label, "LYS", element=label[:1])
    lines.append(_pdb_line("ATOM", 9999, "CB", "ALA", 1.0, 2.0, 3.0, element="C") + " " * 40)  # > 99 chars
    rec("ATOM", "CG", "LEU", element="C")
    lines.append("ENDMDL")
    lines = [ln.ljust(80) for ln in lines]
    tail = [ln.ljust(80) for ln in ("END", _pdb_line("ATOM", 1, "N", "GLY", 0.0, 0.0, 0.0, element="N"))]
    path = tmp_path / "synthetic.pdb"
    path.write_text("\n".join(lines + tail) + "\n")
    for all_models in (True, False):
        b, off = read_pdb(str(path), all_models=all_models)
        rb, roff = po.ref_pdb_read(path, all_models)
        assert np.array_equal(off, roff), all_models
        assert np.array_equal(b.view(np.uint32), rb.view(np.uint32)), all_models
        if all_models:
            assert len(roff) == 4 and roff[2] - roff[1] > 700   # three models, the big one typed
    # no END and no newline after the last record
    path.write_text("\n".join(lines))
    for all_models in (True, False):
        b, off = read_pdb(str(path), all_models=all_models)
        rb, roff = po.ref_pdb_read(path, all_models)
        assert np.array_equal(off, roff) and np.array_equal(b.view(np.uint32), rb.view(np.uint32)), all_models
    assert po.ref_pdb_read(tmp_path / "missing.pdb") == -1100
